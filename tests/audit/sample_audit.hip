// sample_audit.hip -- CPU audit of the beamlet sample stage's indexing.  TEST INFRASTRUCTURE.
//
// Builds the job tables of a workload on the host with the library's own planning code
// (plan_beam, build_job: this file includes endtoend_amd/csrc/beamlet.hip), computes every REV
// line's stored interval and store offset with the interval kernel's own function, and replays
// the sample kernel's passes A.1-B.2 line by line on the CPU with the kernel's own helpers
// (bl_locate, rev_to_packed, bl_terma_point, fluence_cell; point_fluence checks the cell test).  Every index the kernel forms is
// checked against the bound of the array it addresses (and bl_locate, whose quotients skip the
// IEEE division away from multiples, against the division form of the mapping on every call):
//   LDS:    s_fp / s_list (< 1024 and < the line's stored length), s_pil / s_rank words (< 128),
//           staged pillar slots (< kStagePillars), the affine pillar index read in pass B.1;
//   global: the beamlet record (< the job's beamlets), REV store and half-step offsets
//           (< the direction's store size), the trimmed interval (inside the stored one);
//   queue:  requests written in pass B.2 never exceed the line's reservation from pass A.2.
// The pass A.1 cull is checked to keep every footprint sample: on every third audited line, each
// stored sample is classified exactly (mapping, half-width, fluence cell) whether or not the cull
// keeps it, and a footprint sample the cull drops is a violation.
//
//   sample_audit <case dir> [line stride]
// The case directory is the one tests/shim_case.py writes.  Prints one JSON line of counts and
// exits 1 on any violation.
#include "../../endtoend_amd/csrc/beamlet.hip"

#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <random>
#include <sstream>

namespace {

struct Audit {
    uint64_t lines = 0, nonempty = 0, cand_lines = 0, foot_lines = 0, exact_lines = 0;
    uint64_t samples = 0, cand = 0, foot = 0, exact = 0, requests = 0, shortfall = 0, staged = 0,
             stage_overflow = 0, violations = 0, locates = 0, cull_checked = 0;
    uint32_t max_len = 0, max_pil_word = 0, max_stage = 0;
    std::string first;
    void merge(const Audit& o) {
        lines += o.lines; nonempty += o.nonempty; cand_lines += o.cand_lines; foot_lines += o.foot_lines;
        exact_lines += o.exact_lines; samples += o.samples; cand += o.cand; foot += o.foot; exact += o.exact;
        requests += o.requests; shortfall += o.shortfall; staged += o.staged; stage_overflow += o.stage_overflow;
        violations += o.violations; locates += o.locates; cull_checked += o.cull_checked;
        max_len = std::max(max_len, o.max_len);
        max_pil_word = std::max(max_pil_word, o.max_pil_word);
        max_stage = std::max(max_stage, o.max_stage);
        if (first.empty()) first = o.first;
    }
};

#define AUDIT(cond, what)                                                                             \
    do {                                                                                              \
        if (!(cond)) {                                                                                \
            A.violations++;                                                                           \
            if (A.first.empty()) {                                                                    \
                std::ostringstream os;                                                                \
                os << what << " (line " << line << ")";                                               \
                A.first = os.str();                                                                   \
            }                                                                                         \
        }                                                                                             \
    } while (0)

// The pillar mapping with IEEE quotients throughout (pb / pd floored, fmod_pos): what bl_locate
// must equal bit for bit.
bool locate_div(const BlArgs& a, const BlConeDev& c, const BatchDev& B, int tX, int tY, int tZ, int& pidx, v3& inp) {
    const v3 pb = rev_to_packed(c, a.rrv, tX, tY, tZ);
    const v3 gdf = mk3((float)B.gd[0], (float)B.gd[1], (float)B.gd[2]);
    if (pb.x < 0.f || pb.y < 0.f || pb.z < 0.f || pb.x >= gdf.x || pb.y >= gdf.y || pb.z >= gdf.z) return false;
    const v3 pdf = mk3((float)B.pd[0], (float)B.pd[1], (float)B.pd[2]);
    pidx = f2i_rd(pb.x / pdf.x) + f2i_rd(pb.z / pdf.z) * B.npx;
    if (pidx >= B.n) return false;
    inp = mk3(fmod_pos(pb.x, pdf.x), fmod_pos(pb.y, pdf.y), fmod_pos(pb.z, pdf.z));
    if (inp.x < (float)kWall || inp.y < (float)kWall || inp.z < (float)kWall) return false;
    inp = inp - 0.5f;
    return true;
}

bool same_bits(float x, float y) { return std::memcmp(&x, &y, sizeof(float)) == 0; }

// Passes A.1-B.2 of bl_sample_kernel for one line, with its bounds.
void audit_line(const BlArgs& a, const BlConeDev& c, const BatchDev& B, size_t nbeamlets, uint64_t line,
                LineMeta meta, uint64_t store, Audit& A) {
    A.lines++;
    const int m = meta.len;
    if (m <= 0) return;
    A.nonempty++;
    A.samples += (uint64_t)m;
    A.max_len = std::max(A.max_len, (uint32_t)m);
    AUDIT(m <= kMaxLine, "stored length above 1024");
    AUDIT((uint64_t)meta.off + (uint64_t)m <= store, "stored interval beyond the REV store");
    if (m > kMaxLine) return;
    const uint64_t l = line - c.line0;
    const int tY = (int)(l % (uint64_t)c.ny), tZ = (int)(l / (uint64_t)c.ny);
    const int i0 = meta.i0;
    AUDIT(i0 >= 0 && i0 + m <= c.nx, "stored interval outside the REV line");
    const v3 q0 = mk3(fmaf(c.minc.x, a.rrv.x, (float)tZ), c.minc.y * a.rrv.y, fmaf(c.minc.z, a.rrv.z, (float)tY));
    const v3 p0 = c.kinv_m(q0) + 0.5f;
    const v3 dl = mk3(c.kinv_m.m[1], c.kinv_m.m[4], c.kinv_m.m[7]);
    const float pdx = (float)B.pd[0], pdz = (float)B.pd[2], ipdx = 1.f / pdx, ipdz = 1.f / pdz;
    const float cull = B.fw + 0.05f;
    const v3 ipd = mk3(ipdx, 1.f / (float)B.pd[1], ipdz);
    // the kernel's mapping, checked against the division form
    auto locate = [&](int tX, int& pidx, v3& inp) {
        int pd = -1;
        v3 id = mk3(0.f, 0.f, 0.f);
        const bool ok = bl_locate(a, c, B, ipd, tX, tY, tZ, pidx, inp);
        const bool okd = locate_div(a, c, B, tX, tY, tZ, pd, id);
        A.locates++;
        AUDIT(ok == okd && (!ok || (pidx == pd && same_bits(inp.x, id.x) && same_bits(inp.y, id.y) &&
                                    same_bits(inp.z, id.z))),
              "pillar mapping differs from its division form");
        return ok;
    };
    const bool track = B.n <= (int)kPillarBits;
    uint8_t s_fp[kMaxLine];
    short s_list[kMaxLine];
    uint32_t s_pil[kPillarBits / 32] = {};
    // A.1
    int ncand = 0;
    for (int j = 0; j < m; j++) {
        const float t = (float)(i0 + j);
        const float px = p0.x + t * dl.x, pz = p0.z + t * dl.z, py = p0.y + t * dl.y;
        const float xl = px - floorf(px * ipdx) * pdx, zl = pz - floorf(pz * ipdz) * pdz;
        const bool cand = fabsf(xl - B.fc[0]) <= cull && fabsf(zl - B.fc[1]) <= cull &&
                          xl >= B.cb[0] + B.cb[1] * py && xl <= B.cb[2] + B.cb[3] * py &&
                          zl >= B.cb[4] + B.cb[5] * py && zl <= B.cb[6] + B.cb[7] * py;
        s_fp[j] = 0;
        if (line % 3 == 0) {  // every stored sample classified exactly: the cull must keep each footprint sample
            int pe = -1;
            v3 ie;
            if (bl_locate(a, c, B, ipd, i0 + j, tY, tZ, pe, ie) && fabsf(ie.x - B.fc[0]) <= B.fw &&
                fabsf(ie.z - B.fc[1]) <= B.fw && pe >= 0 && (size_t)(B.beamlet0 + pe) < nbeamlets) {
                const PillarRec be = pillar_rec(a.beamlets[B.beamlet0 + pe]);
                if (fluence_cell(bl_terma_point(a, B, be, ie), B.beam) == be.fidx) {
                    A.cull_checked++;
                    AUDIT(cand, "the pass A.1 cull drops a footprint sample");
                }
            }
        }
        if (cand) {
            AUDIT(ncand < kMaxLine, "candidate list beyond 1024");
            s_list[ncand++] = (short)j;
        }
    }
    if (ncand == 0) return;
    A.cand_lines++;
    A.cand += (uint64_t)ncand;
    // A.2
    uint32_t nfoot = 0;
    for (int r = 0; r < ncand; r++) {
        const int j = s_list[r];
        AUDIT(j >= 0 && j < m, "candidate index outside the line");
        int pidx = -1;
        v3 inp;
        if (locate(i0 + j, pidx, inp) && fabsf(inp.x - B.fc[0]) <= B.fw &&
            fabsf(inp.z - B.fc[1]) <= B.fw) {
            AUDIT(pidx >= 0 && pidx < B.n, "pillar index outside the batch");
            AUDIT((size_t)(B.beamlet0 + pidx) < nbeamlets, "beamlet record beyond the job's beamlets");
            if (pidx < 0 || (size_t)(B.beamlet0 + pidx) >= nbeamlets) continue;
            const PillarRec bl = pillar_rec(a.beamlets[B.beamlet0 + pidx]);
            // the kernel's cell test, and the fluence test it stands for (a job beamlet's own cell
            // has a positive weight)
            const v3 tp = bl_terma_point(a, B, bl, inp);
            const bool fp = fluence_cell(tp, B.beam) == bl.fidx;
            AUDIT(fp == (point_fluence(tp, B.fmap, B.beam, bl.fidx) > 0.f), "cell test differs from the fluence test");
            if (fp) {
                s_fp[j] = 1;
                nfoot++;
                if (track) {
                    AUDIT((pidx >> 5) < (int)(kPillarBits / 32), "footprint pillar word beyond 128");
                    A.max_pil_word = std::max(A.max_pil_word, (uint32_t)(pidx >> 5));
                    s_pil[pidx >> 5] |= 1u << (pidx & 31);
                }
            }
        }
    }
    if (nfoot == 0) return;
    A.foot_lines++;
    A.foot += nfoot;
    // staging: ranks and slots
    uint32_t s_rank[kPillarBits / 32];
    uint32_t nstage = 0;
    if (track) {
        for (int w = 0; w < (int)(kPillarBits / 32); w++) {
            s_rank[w] = nstage;
            nstage += (uint32_t)__builtin_popcount(s_pil[w]);
        }
        A.max_stage = std::max(A.max_stage, nstage);
        A.staged += std::min<uint32_t>(nstage, kStagePillars);
        if (nstage > (uint32_t)kStagePillars) A.stage_overflow++;
    }
    // B.1
    int nex = 0;
    for (int j = 0; j < m; j++) {
        bool exact = true;
        if (track) {
            const float t = (float)(i0 + j);
            const float px = p0.x + t * dl.x, pz = p0.z + t * dl.z;
            const float fx = floorf(px * ipdx), fz = floorf(pz * ipdz);
            const float xl = px - fx * pdx, zl = pz - fz * pdz;
            if (xl > 0.05f && xl < pdx - 0.05f && zl > 0.05f && zl < pdz - 0.05f) {
                const int pa = (int)fx + (int)fz * B.npx;
                const bool outside = fx < 0.f || fz < 0.f || (int)fx >= B.npx || pa >= B.n;
                if (!outside) AUDIT(pa >= 0 && (pa >> 5) < (int)(kPillarBits / 32), "affine pillar word beyond 128");
                exact = !(outside || !((s_pil[pa >> 5] >> (pa & 31)) & 1u));
            }
        }
        if (exact) {
            s_fp[j] |= 2;
            s_list[nex++] = (short)j;
        }
    }
    if (nex == 0) return;
    A.exact_lines++;
    A.exact += (uint64_t)nex;
    const int elo = s_list[0], ehi = s_list[nex - 1];
    AUDIT(elo >= 0 && ehi < m && elo <= ehi, "wall span outside the line");
    // B.2
    uint32_t qdone = 0;
    int nlo = m, nhi = -1;
    for (int r = 0; r < nex; r++) {
        const int j = s_list[r];
        AUDIT((uint64_t)meta.off + (uint64_t)j < store, "half step / TERMA slot beyond the REV store");
        int pidx = -1;
        v3 inp;
        bool h = false;
        if (locate(i0 + j, pidx, inp) && (!track || ((s_pil[pidx >> 5] >> (pidx & 31)) & 1u))) {
            AUDIT(pidx >= 0 && pidx < B.n, "pillar index outside the batch");
            if (track) {
                const uint32_t slot = s_rank[pidx >> 5] + (uint32_t)__builtin_popcount(s_pil[pidx >> 5] & ((1u << (pidx & 31)) - 1u));
                AUDIT(slot < nstage, "staged slot beyond the footprint pillars");
            }
            h = true;
            if (s_fp[j] & 1) qdone++;
        }
        if (h) {
            nlo = std::min(nlo, j);
            nhi = std::max(nhi, j);
        }
    }
    AUDIT(qdone <= nfoot, "TERMA requests beyond the line's reservation");
    A.requests += qdone;
    A.shortfall += nfoot - std::min(qdone, nfoot);
    if (nhi >= nlo) AUDIT(nlo >= 0 && nhi < m, "trimmed interval outside the stored one");
}

// floor_quot against the floor of the IEEE quotient: every divisor 1..256 (pillar dimensions are
// small integers), 13 consecutive floats around each of the first 4096 multiples of it, and random
// dividends below 2^14 times it.
void check_floor_quot(uint64_t& checked, uint64_t& bad) {
    std::mt19937_64 g(20261202);
    for (int yi = 1; yi <= 256; yi++) {
        const float y = (float)yi, iy = 1.f / y;
        auto one = [&](float x) {
            checked++;
            if (!same_bits(floor_quot(x, y, iy), floorf(x / y))) bad++;
        };
        for (int k = 0; k <= 4096; k++) {
            float x = (float)k * y;
            for (int t = 0; t < 6 && x > 0.f; t++) x = std::nextafter(x, 0.f);
            for (int t = 0; t < 13; t++, x = std::nextafter(x, 1e30f)) one(x);
        }
        std::uniform_real_distribution<float> U(0.f, 16384.f * y);
        for (int t = 0; t < 20000; t++) one(U(g));
    }
}

std::vector<float> read_raw(const std::string& path, size_t n) {
    std::vector<float> v(n);
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(float))))
        throw std::runtime_error("cannot read " + path);
    return v;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <case dir> [line stride]\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const uint64_t stride = argc > 2 ? std::stoull(argv[2]) : 1;
    std::ifstream cf(dir + "/case.txt");
    std::map<std::string, std::vector<double>> kv;
    std::vector<std::vector<double>> bls;
    for (std::string ln; std::getline(cf, ln);) {
        std::istringstream is(ln);
        std::string key;
        if (!(is >> key)) continue;
        std::vector<double> vals;
        for (double x; is >> x;) vals.push_back(x);
        if (key == "beam") bls.push_back(vals);
        else kv[key] = vals;
    }
    cccs_ctx ctx;
    cccs_problem& p = ctx.prob;
    for (int k = 0; k < 3; k++) {
        p.size[k] = (uint32_t)kv.at("size")[k];
        p.voxel[k] = (float)kv.at("voxel")[k];
        p.start[k] = (float)kv.at("start")[k];
        p.bbox_start[k] = (uint32_t)kv.at("bbox_start")[k];
        p.bbox_size[k] = (uint32_t)kv.at("bbox_size")[k];
        p.max_rev_size[k] = (uint32_t)kv.at("max_rev_size")[k];
    }
    p.rev_latspacing = (float)kv.at("rev_latspacing")[0];
    p.rev_longspacing = (float)kv.at("rev_longspacing")[0];
    p.kernel_extent = (float)kv.at("kernel_extent")[0];
    p.penumbra = (float)kv.at("penumbra")[0];
    p.ss_factor = (uint32_t)kv.at("ss_factor")[0];
    p.nphi = (uint32_t)kv.at("nphi")[0];
    p.ntheta = (uint32_t)kv.at("ntheta")[0];
    p.nradii = (uint32_t)kv.at("nradii")[0];
    ctx.nrays = (int)(p.nphi * (p.ntheta / 2));
    ctx.theta = read_raw(dir + "/theta", (size_t)ctx.nrays);
    ctx.phi = read_raw(dir + "/phi", (size_t)ctx.nrays);
    std::vector<cccs_beam> beams(bls.size());
    std::vector<std::vector<float>> fmaps(bls.size());
    for (size_t i = 0; i < bls.size(); i++) {
        const std::vector<double>& b = bls[i];
        cccs_beam& o = beams[i];
        o.azimuth = (float)b[1]; o.zenith = (float)b[2]; o.coll = (float)b[3];
        for (int k = 0; k < 3; k++) {
            o.source[k] = (float)b[4 + k];
            o.direction[k] = (float)b[7 + k];
            o.isocenter[k] = (float)b[10 + k];
        }
        o.sad = (float)b[13];
        o.fmap_size[0] = (uint32_t)b[14]; o.fmap_size[1] = (uint32_t)b[15];
        o.beamlet_size[0] = (float)b[16]; o.beamlet_size[1] = (float)b[17];
        fmaps[i] = read_raw(dir + "/fmap_" + std::to_string(i), (size_t)o.fmap_size[0] * o.fmap_size[1]);
    }
    std::vector<BeamPlan> plans;
    for (size_t i = 0; i < beams.size(); i++) {
        BeamPlan P;
        if (plan_beam(&ctx, beams[i], fmaps[i].data(), (int)i, P) != CCCS_OK) {
            std::fprintf(stderr, "plan: %s\n", cccs_last_error());
            return 2;
        }
        if (!P.bl.empty()) plans.push_back(std::move(P));
    }
    std::vector<const float*> host_fmaps;
    for (const BeamPlan& P : plans) host_fmaps.push_back(fmaps[(size_t)P.beam].data());
    const std::vector<BatchRef> refs = all_batches(plans);
    Job J;
    if (build_job(&ctx, plans, refs, host_fmaps, J) != CCCS_OK) {
        std::fprintf(stderr, "job: %s\n", cccs_last_error());
        return 2;
    }
    BlArgs a = base_args(&ctx);
    a.batches = J.batches.data();
    a.beamlets = J.beamlets.data();
    a.nbatch = (int)J.batches.size();
    const int nb = (int)J.batches.size();
    Audit total;
    for (int r = 0; r < ctx.nrays; r++) {
        a.cones = J.cones.data() + (size_t)r * nb;
        const uint64_t nl = J.ray_lines[r];
        // interval pass and the exclusive scan of the stored lengths
        std::vector<LineMeta> meta(nl);
        uint64_t store = 0;
        for (uint64_t line = 0; line < nl; line++) {
            const int ci = find_cone(a.cones, a.nbatch, line);
            meta[line] = line_interval(a, a.cones[ci], a.batches[ci], line);
            meta[line].off = (uint32_t)store;
            store += (uint64_t)std::max<int>(0, meta[line].len);
        }
#pragma omp parallel
        {
            Audit A;
#pragma omp for schedule(dynamic, 256)
            for (uint64_t line = 0; line < nl; line += stride) {
                const int ci = find_cone(a.cones, a.nbatch, line);
                audit_line(a, a.cones[ci], a.batches[ci], J.beamlets.size(), line, meta[line], store, A);
            }
#pragma omp critical
            total.merge(A);
        }
    }
    uint64_t fq_checked = 0, fq_bad = 0;
    check_floor_quot(fq_checked, fq_bad);
    total.violations += fq_bad;
    if (fq_bad && total.first.empty()) total.first = "floor_quot differs from the IEEE quotient's floor";
    std::printf("{\"beams\": %zu, \"batches\": %d, \"beamlets\": %zu, \"directions\": %d, \"lines\": %llu, "
                "\"nonempty\": %llu, \"candidate_lines\": %llu, \"footprint_lines\": %llu, \"exact_lines\": %llu, "
                "\"samples\": %llu, \"candidates\": %llu, \"footprint\": %llu, \"exact\": %llu, \"requests\": %llu, "
                "\"shortfall\": %llu, \"staged\": %llu, \"stage_overflow_lines\": %llu, \"max_len\": %u, "
                "\"max_pillar_word\": %u, \"max_footprint_pillars\": %u, \"locates\": %llu, \"cull_checked\": %llu, "
                "\"floor_quot_checked\": %llu, \"floor_quot_mismatch\": %llu, \"violations\": %llu, \"first\": \"%s\"}\n",
                plans.size(), nb, J.beamlets.size(), ctx.nrays, (unsigned long long)total.lines,
                (unsigned long long)total.nonempty, (unsigned long long)total.cand_lines,
                (unsigned long long)total.foot_lines, (unsigned long long)total.exact_lines,
                (unsigned long long)total.samples, (unsigned long long)total.cand, (unsigned long long)total.foot,
                (unsigned long long)total.exact, (unsigned long long)total.requests,
                (unsigned long long)total.shortfall, (unsigned long long)total.staged,
                (unsigned long long)total.stage_overflow, total.max_len, total.max_pil_word, total.max_stage,
                (unsigned long long)total.locates, (unsigned long long)total.cull_checked,
                (unsigned long long)fq_checked, (unsigned long long)fq_bad,
                (unsigned long long)total.violations, total.first.c_str());
    return total.violations ? 1 : 0;
}
