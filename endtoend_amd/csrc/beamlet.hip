// beamlet.hip -- beamlet dose coefficients (dosecalc-beamlet: nvbbRayConvolve_host.cu:444-1409,
// nvbbRayConvolve_device.cu:19-521) on gfx950.
//
// Every active beamlet of a beam becomes a pillar (its beamlet's path through the calc_bbox plus a
// kernel_extent margin and a wall), and the pillars of a batch are tiled side by side in one packed
// BEV grid.  For each cone direction, every REV line of every batch in the job is resampled into
// LDS (pillar mapping, per-beamlet TERMA, density), walked inside its pillar segment, and the cone
// dose is gathered into the packed grid.  Each beamlet's pillar is then resampled back onto the
// calc_bbox and compacted on the device into a sparse column.
//
// A job is a list of batches, possibly from several beams: one launch of each kernel covers every
// batch of the job, so small beams and large beams share launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <thread>
#include <vector>

#include "context.hpp"
#include "device_common.hpp"

namespace cccs {

constexpr int kWall = 2;        // pillar wall, REV samples (PILLAR_GRID, io_data_structs.h:139-160)
constexpr int kBuffer = 2;      // pillar rim excluded from unpacking (same struct)
constexpr int kMaxBatches = 100;  // batch-count retry limit of the reference (host.cu:692-834)
constexpr uint32_t kSubQueues = 64;  // TERMA sub-queues per cone direction
constexpr uint32_t kCntStride = 32;  // a sub-queue counter per 128-byte line: no two share one
constexpr uint32_t kPillarBits = 4096;  // batches up to this many pillars track footprint pillars per line

// One Rodrigues rotation with host-evaluated sine/cosine.
struct Rot1 {
    v3 a;
    float s, c;
};
CCCS_HD v3 rot1(const Rot1& R, v3 p) { return rodrigues(p, R.a, R.s, R.c); }

// invRotateBeamlet / rotateBeamlet (CudaUtilities/geometry.cu:99-115): a rotation about
// (sin(-phi), 0, cos(-phi)) by -theta (beamlet frame -> BEV) or +theta (BEV -> beamlet frame).
static Rot1 beamlet_rot(float theta, float phi, bool to_bev) {
    Rot1 r;
    float s, c;
    sincosf(-phi, &s, &c);
    r.a = mk3(s, 0.f, c);
    sincosf(to_bev ? -theta : theta, &r.s, &r.c);
    return r;
}

struct BeamletDev {
    v3 pst;          // pillar origin [cm]
    Rot1 inv;        // beamlet frame -> BEV
    Rot1 fwd;        // BEV -> beamlet frame
    int fidx;        // fluence-map index of the beamlet
    int batch;       // batch of the job
    int pix, piy;    // pillar column / row inside the batch grid
    int box0[3];     // calc_bbox-relative voxel box that can receive its dose
    int boxn[3];
    uint64_t arena;  // first element of its box in the column arena (i outer, k inner)
    uint32_t row0;   // first of its box's (i, j) rows in the job's row counts
    m33 fwd_m;       // PVCS -> beamlet frame (beamlet forward * beam forward), matrix form
};

struct BatchDev {
    BeamDev beam;
    Rot2 beam_inv;         // BEV -> PVCS
    const float* fmap;     // fluence map of the beam (device)
    int n, npx, npy;       // beamlets, pillars per row, rows
    int gd[3];             // packed grid dims (pX, pY, pZ)
    int pd[3];             // pillar dims
    int beamlet0;          // first beamlet of the batch in the job's beamlet table
    uint64_t pdose;        // float offset of its packed dose grid ([pX][pZ][pY], pY fastest)
    uint32_t tile0;        // first gather tile
    int tiles[3];          // gather tiles along pY (16), pZ (4), pX (4)
    float fc[2];           // beamlet axis in pillar sample coordinates (x, z)
    float fw;              // bound on the footprint half-width [samples]: farther samples get no TERMA
    float cb[8];           // footprint bounds linear in depth (footprint_cull): x lo, x hi, z lo, z hi
    float gw;              // largest footprint distance from the axis over the grid's depth (<= fw)
};

// A footprint sample whose TERMA needs a Siddon march.
constexpr uint32_t kNoSlot = 0xffffffffu;  // a reserved queue entry that carries no request

struct TermaReq {
    uint32_t slot;     // position in the REV store of the cone direction, or kNoSlot
    uint32_t beamlet;  // job beamlet index
    float x, y, z;     // TERMA point [cm]
};

// One (cone, batch) REV grid.
struct BlConeDev {
    Rot2 kinv;       // REV -> packed BEV (invRotateKernel)
    Rot2 kfwd;       // packed BEV -> REV (rotateKernel)
    m33 kinv_m;      // matrix form of kinv, for line intervals only
    m33 kfwd_m;      // matrix form of kfwd, for the packed-BEV gather
    v3 minc;         // REV grid origin [cm]
    int nx, ny, nz;  // line length, lines along y, lines along z
    int batch;
    uint64_t line0;  // first line of this grid among the launch's lines
};

struct BlArgs {
    const float* dens;
    const float* tables;        // kernel rows | radii | 1/width | mean radii | 1/width of mean radii
    const BatchDev* batches;
    const BlConeDev* cones;     // the grids of one cone direction, one per batch
    const BeamletDev* beamlets;
    LineMeta* lines;
    float* rev;
    float* pdose;
    int nbatch;
    uint64_t nlines;
    int ntheta, nradii, nphi, rowf, rowb;
    float delr;
    v3 rvs, rrv;                // REV spacings (lat, long, lat) and reciprocals
    VolDesc vd;
    SpecDev spec;
    float bhc;
    uint32_t* line_cone;        // grid (= batch) of every line, written by the interval pass
    uint32_t* counts;
    const uint32_t* tile_batch; // batch of every gather tile (nullptr: a single batch)
};

CCCS_HD int find_cone(const BlConeDev* __restrict__ c, int n, uint64_t line) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (c[mid].line0 <= line) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// REV sample (tX, tY, tZ) of a grid -> packed-BEV coordinate (cudaBeamletRaytrace,
// nvbbRayConvolve_device.cu:33-41).
CCCS_HD v3 rev_to_packed(const BlConeDev& c, v3 rrv, int tX, int tY, int tZ) {
    v3 q = mk3(fmaf(c.minc.x, rrv.x, (float)tZ), fmaf(c.minc.y, rrv.y, (float)tX), fmaf(c.minc.z, rrv.z, (float)tY));
    return rot2(c.kinv, q) + 0.5f;
}

// Interval of stored samples of every line of one cone direction: the line clipped to the packed
// grid box (packed coordinates are affine in the sample index), widened by 2/3 samples.
CCCS_HD LineMeta line_interval(const BlArgs& a, const BlConeDev& c, const BatchDev& B, uint64_t line) {
    const uint64_t l = line - c.line0;
    const int tY = (int)(l % (uint64_t)c.ny), tZ = (int)(l / (uint64_t)c.ny);
    v3 q0 = mk3(fmaf(c.minc.x, a.rrv.x, (float)tZ), c.minc.y * a.rrv.y, fmaf(c.minc.z, a.rrv.z, (float)tY));
    v3 p0 = c.kinv_m(q0) + 0.5f;
    v3 d = mk3(c.kinv_m.m[1], c.kinv_m.m[4], c.kinv_m.m[7]);
    float t0 = 0.f, t1 = (float)c.nx;
    clip_slab(p0.x, d.x, 0.f, (float)B.gd[0], t0, t1);
    clip_slab(p0.y, d.y, 0.f, (float)B.gd[1], t0, t1);
    clip_slab(p0.z, d.z, 0.f, (float)B.gd[2], t0, t1);
    LineMeta m;
    m.i0 = 0;
    m.len = 0;
    m.off = 0;
    if (t0 <= t1) {
        const int i0 = max(0, (int)floorf(t0) - 2);
        const int i1 = min(c.nx, (int)ceilf(t1) + 3);
        m.i0 = (int16_t)i0;
        m.len = (int16_t)max(0, i1 - i0);
    }
    return m;
}

__global__ void __launch_bounds__(256) bl_interval_kernel(BlArgs a) {
    const uint64_t line = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (line >= a.nlines) return;
    const int ci = find_cone(a.cones, a.nbatch, line);
    a.line_cone[line] = (uint32_t)ci;
    a.lines[line] = line_interval(a, a.cones[ci], a.batches[ci], line);
}

// fmodf(x, y) for 0 <= x < 2^24 and a positive integer y <= 2^20: n*y is exact, x - n*y is exact,
// and the fix-ups absorb a quotient rounded across an integer.  Equal to fmodf bit for bit there.
CCCS_HD float fmod_pos(float x, float y) {
    const float n = truncf(x / y);
    float r = fmaf(-n, y, x);
    if (r < 0.f) r += y;
    if (r >= y) r -= y;
    return r;
}

// floorf(x / y) of the correctly rounded quotient, for 0 <= x < 2^24 and a positive integer
// y <= 2^20, with iy = 1 / y rounded.  q = x * iy lies within q * 2^-22 of the rounded quotient (two
// roundings of 2^-24 relative, plus half an ulp of the quotient), so the two have the same floor
// unless q lies within that distance of an integer; there the division decides.  Lanes near a
// multiple of y are rare, so most waves skip the IEEE division.
CCCS_HD float floor_quot(float x, float y, float iy) {
    const float q = x * iy;
    const float fq = floorf(q);
    const float m = q * 4.76837158e-7f;  // 2^-21: twice the bound
    if (q - fq > m && (fq + 1.f) - q > m) return fq;
    return floorf(x / y);
}

// fmod_pos(x, y) given the floor n of its quotient (the same n fmod_pos computes).
CCCS_HD float fmod_pos_n(float x, float y, float n) {
    float r = fmaf(-n, y, x);
    if (r < 0.f) r += y;
    if (r >= y) r -= y;
    return r;
}

// Pillar mapping of one REV sample (cudaBeamletRaytrace, nvbbRayConvolve_device.cu:33-68):
// false for a wall, a sample outside the packed grid or past the last pillar; otherwise the
// beamlet (pidx, batch-local) and the pillar-local coordinate inp (already shifted by -0.5).
// ipd holds the reciprocals of the pillar dimensions; the pillar index and the in-pillar
// coordinate are those of the IEEE quotients pb / pd (floor_quot).
CCCS_HD bool bl_locate(const BlArgs& a, const BlConeDev& c, const BatchDev& B, v3 ipd, int tX, int tY, int tZ,
                       int& pidx, v3& inp) {
    const v3 pb = rev_to_packed(c, a.rrv, tX, tY, tZ);
    const v3 gdf = mk3((float)B.gd[0], (float)B.gd[1], (float)B.gd[2]);
    if (pb.x < 0.f || pb.y < 0.f || pb.z < 0.f || pb.x >= gdf.x || pb.y >= gdf.y || pb.z >= gdf.z) return false;
    const v3 pdf = mk3((float)B.pd[0], (float)B.pd[1], (float)B.pd[2]);
    const float nx = floor_quot(pb.x, pdf.x, ipd.x), ny = floor_quot(pb.y, pdf.y, ipd.y),
                nz = floor_quot(pb.z, pdf.z, ipd.z);
    pidx = (int)nx + (int)nz * B.npx;
    if (pidx >= B.n) return false;
    inp = mk3(fmod_pos_n(pb.x, pdf.x, nx), fmod_pos_n(pb.y, pdf.y, ny), fmod_pos_n(pb.z, pdf.z, nz));
    if (inp.x < (float)kWall || inp.y < (float)kWall || inp.z < (float)kWall) return false;
    inp = inp - 0.5f;
    return true;
}

// What a REV sample's density and TERMA points need of its beamlet (the head of BeamletDev).
struct PillarRec {
    v3 pst;
    Rot1 inv;
    int fidx;
};
CCCS_HD PillarRec pillar_rec(const BeamletDev& d) {
    PillarRec r;
    r.pst = d.pst;
    r.inv = d.inv;
    r.fidx = d.fidx;
    return r;
}

// PVCS point [cm] where the sample's TERMA is evaluated (device.cu:170-180).
CCCS_HD v3 bl_terma_point(const BlArgs& a, const BatchDev& B, const PillarRec& bl, v3 inp) {
    const v3 g = rot2(B.beam_inv, rot1(bl.inv, inp));
    return mk3(fmaf(g.x, a.rvs.x, bl.pst.x), fmaf(g.y, a.rvs.y, bl.pst.y), fmaf(g.z, a.rvs.z, bl.pst.z));
}

// Density at the sample (device.cu:213-223): trilinear, border 0.
__device__ __forceinline__ float bl_density(const BlArgs& a, const BatchDev& B, const PillarRec& bl, v3 inp) {
    const v3 g2 = rot2(B.beam_inv, rot1(bl.inv, inp * a.rvs)) + bl.pst;
    const v3 gi = (g2 - a.vd.start) / a.vd.voxel;
    return tex3_border(a.dens, a.vd.nx, a.vd.ny, a.vd.nz, gi.x + 0.5f, gi.y + 0.5f, gi.z + 0.5f);
}

// Cumulative kernel of one row, linear between radii (the reference's clamped linear texture
// lookup, nvbbRayConvolve_device.cu:330-352).  Interval s of a row is {lo, K(lo), slope, hi}: the
// value at r in (lo, hi] is K(lo) + (r - lo) * slope.  s tracks the first radius >= r and r must
// not decrease between calls.
__device__ __forceinline__ float kc_at(float r, const float4* __restrict__ tab, int nr, int& s, float4& iv) {
    while (s < nr - 1 && iv.w < r) iv = tab[++s];
    return iv.y + (r - iv.x) * iv.z;
}

// The grid and batch fields the mapping reads sit in LDS; this compiler-only barrier keeps their
// reads inside each pass iteration (next to their use) instead of holding ~50 of them in VGPRs
// across the pass.
__device__ __forceinline__ void lds_reload_fence() { __asm__ volatile("" ::: "memory"); }

constexpr int kStagePillars = 64;  // footprint pillars of one line whose beamlet data sits in LDS

// Stage 1 of a cone direction, one wave per REV line: pillar mapping of every stored sample and
// the beamlet-footprint test (fluence cell of the sample == its beamlet).  A line without footprint
// samples carries no TERMA, so all its cone dose is zero: it is marked empty.  Otherwise the half
// steps (0.5 * delr * density, -1 for walls) go to H, TERMA slots are zeroed, and the footprint
// samples are queued for their Siddon TERMA.
//
// The per-sample work is a chain of dependent memory trips (line record, grid tables, beamlet
// record, density taps), so the kernel keeps each chain short: the line's grid comes from the line
// table built by the interval pass (no search), its grid and batch tables load together, and the
// beamlet data of the line's footprint pillars is staged in LDS once, while the pillar
// classification runs, instead of being read per sample by the density pass.
__global__ void __launch_bounds__(64) bl_sample_kernel(BlArgs a, float* __restrict__ H, TermaReq* __restrict__ q,
                                                       uint32_t subcap, uint32_t* __restrict__ qcnt) {
    __shared__ uint8_t s_fp[kMaxLine];            // bit 0: footprint sample, bit 1: on the exact list
    __shared__ short s_list[kMaxLine];            // compacted work list of the current pass
    __shared__ uint32_t s_pil[kPillarBits / 32];  // pillars with footprint samples on this line
    __shared__ uint16_t s_rank[kPillarBits / 32]; // footprint pillars in the words before each word
    __shared__ PillarRec s_rec[kStagePillars];    // their beamlet data, in pillar order
    __shared__ BlConeDev s_cone;                  // the line's grid and batch records
    __shared__ BatchDev s_batch;
    static_assert(sizeof(BlConeDev) % 8 == 0 && sizeof(BatchDev) % 8 == 0 &&
                      (sizeof(BlConeDev) + sizeof(BatchDev)) / 8 <= 64, "one 8-byte word per lane");
    uint32_t stored = 0;
    uint32_t staged = 0xffffffffu;  // grid whose records sit in LDS
    for (uint64_t line = xcd_remap(blockIdx.x, gridDim.x); line < a.nlines; line += gridDim.x) {
    const LineMeta meta = a.lines[line];
    const uint32_t ci = a.line_cone[line];
    if (meta.len <= 0) continue;
    // the lane id is opaque per line, so that the lane arithmetic is redone per line rather than
    // hoisted out of the line loop into registers held for the whole kernel
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    const uint64_t lt = (1ull << tid) - 1ull;
    // each wave walks a stride of lines; a line's TERMA requests go to one of kSubQueues queues,
    // chosen by its run of 64 consecutive lines: the Siddon stage then marches neighbouring
    // lines' rays side by side (shared cache lines; -26 % stage time against one queue per wave),
    // and the lines in flight still spread over every counter
    const uint32_t sq = (uint32_t)((line >> 6) % kSubQueues);
    TermaReq* qs = q + (size_t)sq * subcap;
    // grid ci of a direction belongs to batch ci; both records go to LDS when the line's grid
    // changes, so the passes below read their fields there instead of from global memory
    if (ci != staged) {
        constexpr int nc = (int)(sizeof(BlConeDev) / 8), nb = (int)(sizeof(BatchDev) / 8);
        if (tid < nc) reinterpret_cast<uint64_t*>(&s_cone)[tid] = reinterpret_cast<const uint64_t*>(a.cones + ci)[tid];
        else if (tid < nc + nb)
            reinterpret_cast<uint64_t*>(&s_batch)[tid - nc] = reinterpret_cast<const uint64_t*>(a.batches + ci)[tid - nc];
        staged = ci;
        __syncthreads();
    }
    const BlConeDev& c = s_cone;
    const BatchDev& B = s_batch;
    const uint64_t l = line - c.line0;
    const int tY = (int)(l % (uint64_t)c.ny), tZ = (int)(l / (uint64_t)c.ny);
    const int i0 = meta.i0, m = meta.len;
    if (tid == 0) stored += (uint32_t)m;  // summed over the wave's lanes by wave_count
    // affine form of the packed coordinate along the line (matrix rotation): used only to skip
    // samples that are far from every footprint or inside pillars without footprint samples;
    // the exact mapping decides every other sample.  Pass A.1 keeps a sample only inside the
    // footprint half-width fw and inside the batch's depth-linear footprint bounds (cb).
    const v3 q0 = mk3(fmaf(c.minc.x, a.rrv.x, (float)tZ), c.minc.y * a.rrv.y, fmaf(c.minc.z, a.rrv.z, (float)tY));
    const v3 p0 = c.kinv_m(q0) + 0.5f;
    const v3 dl = mk3(c.kinv_m.m[1], c.kinv_m.m[4], c.kinv_m.m[7]);
    const float pdx = (float)B.pd[0], pdz = (float)B.pd[2], ipdx = 1.f / pdx, ipdz = 1.f / pdz;
    const float cull = B.fw + 0.05f;
    const v3 ipd = mk3(ipdx, 1.f / (float)B.pd[1], ipdz);
    const bool track = B.n <= (int)kPillarBits;
    for (int k = tid; k < (int)(kPillarBits / 32); k += 64) s_pil[k] = 0u;
    // pass A.1: affine cull of every sample, candidates compacted into s_list
    int ncand = 0;
    for (int base = 0; base < m; base += 64) {
        const int j = base + tid;
        bool cand = false;
        if (j < m) {
            const float t = (float)(i0 + j);
            const float px = p0.x + t * dl.x, pz = p0.z + t * dl.z, py = p0.y + t * dl.y;
            const float xl = px - floorf(px * ipdx) * pdx, zl = pz - floorf(pz * ipdz) * pdz;
            cand = fabsf(xl - B.fc[0]) <= cull && fabsf(zl - B.fc[1]) <= cull && xl >= B.cb[0] + B.cb[1] * py &&
                   xl <= B.cb[2] + B.cb[3] * py && zl >= B.cb[4] + B.cb[5] * py && zl <= B.cb[6] + B.cb[7] * py;
            s_fp[j] = 0;
        }
        const uint64_t mask = __ballot(cand);
        if (cand) s_list[ncand + __popcll(mask & lt)] = (short)j;
        ncand += __popcll(mask);
    }
    if (ncand == 0) {
        if (tid == 0) a.lines[line].len = 0;
        __syncthreads();
        continue;
    }
    __syncthreads();
    // pass A.2: exact pillar mapping and fluence-cell test of the candidates, full waves
    bool any = false;
    uint32_t nfoot = 0;  // footprint samples of the line: exactly its TERMA requests below
    for (int r = 0; r < ncand; r += 64) {
        lds_reload_fence();
        bool fp = false;
        if (r + tid < ncand) {
            const int j = s_list[r + tid];
            int pidx;
            v3 inp;
            if (bl_locate(a, c, B, ipd, i0 + j, tY, tZ, pidx, inp) && fabsf(inp.x - B.fc[0]) <= B.fw &&
                fabsf(inp.z - B.fc[1]) <= B.fw) {
                // point_fluence(...) > 0 without its map load: a job beamlet's own cell has a
                // positive weight (plan_beam keeps only those), so the cell test decides alone
                const PillarRec bl = pillar_rec(a.beamlets[B.beamlet0 + pidx]);
                const v3 tp = bl_terma_point(a, B, bl, inp);
                lds_reload_fence();
                fp = fluence_cell(tp, B.beam) == bl.fidx;
                if (fp) {
                    s_fp[j] = 1;
                    if (track) atomicOr(&s_pil[pidx >> 5], 1u << (pidx & 31));
                }
            }
        }
        const uint64_t fm = __ballot(fp);
        any = any || fm != 0ull;
        nfoot += (uint32_t)__popcll(fm);
    }
    if (!any) {
        if (tid == 0) a.lines[line].len = 0;
        __syncthreads();
        continue;
    }
    // one reservation per line in the sub-queue (its latency overlaps pass B.1): every footprint
    // sample, and no other, queues one TERMA request in pass B.2
    uint32_t qres = 0;
    if (tid == 0) qres = atomicAdd(qcnt + sq * kCntStride, nfoot);
    __syncthreads();
    // the beamlet data of the footprint pillars into LDS (its loads overlap pass B.1): word w of
    // s_pil is lane w's (w < 64) or lane w-64's; a pillar's slot is its rank among the set bits
    if (track) {
        const uint32_t w0 = s_pil[tid], w1 = s_pil[tid + 64];
        const uint32_t c0 = (uint32_t)__popc(w0), c1 = (uint32_t)__popc(w1);
        uint32_t i0s = c0, i1s = c1;  // inclusive scans over the lanes
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u0 = __shfl_up(i0s, o, 64), u1 = __shfl_up(i1s, o, 64);
            if (tid >= o) { i0s += u0; i1s += u1; }
        }
        const uint32_t tot0 = __shfl(i0s, 63, 64);
        const uint32_t r0 = i0s - c0, r1 = tot0 + i1s - c1;
        s_rank[tid] = (uint16_t)r0;
        s_rank[tid + 64] = (uint16_t)r1;
        uint32_t w = w0, slot = r0;
        for (int h = 0; h < 2; h++) {
            const int wbase = (h == 0 ? tid : tid + 64) * 32;
            for (; w != 0u && slot < (uint32_t)kStagePillars; slot++) {
                const int b = __ffs(w) - 1;
                w &= w - 1u;
                s_rec[slot] = pillar_rec(a.beamlets[B.beamlet0 + wbase + b]);
            }
            w = w1;
            slot = r1;
        }
    }
    // pass B.1: a pillar without footprint samples on this line gets no dose on it, so its samples
    // act as walls (a pillar crossing is exactly one segment of the walk).  Samples confidently
    // in such pillars (or outside the grid's pillars) are walls; the others go to the exact list
    // (and are marked with bit 1 of s_fp).  Only the walls between the first and the last listed
    // sample are written: the line is trimmed to its non-wall span below, so no stage reads the
    // others.
    float* T = a.rev + meta.off;
    float* Hl = H + meta.off;
    int nex = 0;
    for (int base = 0; base < m; base += 64) {
        const int j = base + tid;
        bool exact = false;
        if (j < m) {
            exact = true;
            if (track) {
                const float t = (float)(i0 + j);
                const float px = p0.x + t * dl.x, pz = p0.z + t * dl.z;
                const float fx = floorf(px * ipdx), fz = floorf(pz * ipdz);
                const float xl = px - fx * pdx, zl = pz - fz * pdz;
                if (xl > 0.05f && xl < pdx - 0.05f && zl > 0.05f && zl < pdz - 0.05f) {
                    const int pa = (int)fx + (int)fz * B.npx;
                    exact = !(fx < 0.f || fz < 0.f || (int)fx >= B.npx || pa >= B.n ||
                              !((s_pil[pa >> 5] >> (pa & 31)) & 1u));
                }
            }
            if (exact) s_fp[j] |= 2;
        }
        const uint64_t mask = __ballot(exact);
        if (exact) s_list[nex + __popcll(mask & lt)] = (short)j;
        nex += __popcll(mask);
    }
    const uint32_t qbase = __shfl(qres, 0, 64);
    if (nex == 0) {  // every sample is a wall: no dose on this line, and the reservation stays empty
        if (tid == 0) a.lines[line].len = 0;
        for (uint32_t k = (uint32_t)tid; k < nfoot; k += 64)
            if (qbase + k < subcap) qs[qbase + k].slot = kNoSlot;
        __syncthreads();
        continue;
    }
    __syncthreads();
    {
        const int elo = s_list[0], ehi = s_list[nex - 1];
        for (int j = elo + tid; j <= ehi; j += 64)
            if (!(s_fp[j] & 2)) {
                Hl[j] = -1.f;
                T[j] = 0.f;
            }
    }
    // pass B.2: exact mapping, density and TERMA request of the listed samples, full waves
    int nlo = m, nhi = -1;  // first and last sample that is not a wall
    uint32_t qdone = 0;
    for (int r = 0; r < nex; r += 64) {
        lds_reload_fence();
        bool ok = false, need = false;
        int j = 0, pidx = 0;
        v3 inp;
        PillarRec bl;
        if (r + tid < nex) {
            j = s_list[r + tid];
            if (bl_locate(a, c, B, ipd, i0 + j, tY, tZ, pidx, inp) &&
                (!track || ((s_pil[pidx >> 5] >> (pidx & 31)) & 1u))) {
                ok = true;
                const uint32_t slot = track ? (uint32_t)s_rank[pidx >> 5] +
                                                  (uint32_t)__popc(s_pil[pidx >> 5] & ((1u << (pidx & 31)) - 1u))
                                            : (uint32_t)kStagePillars;
                if (slot < (uint32_t)kStagePillars) bl = s_rec[slot];
                else bl = pillar_rec(a.beamlets[B.beamlet0 + pidx]);
                need = (s_fp[j] & 1) != 0;
            }
        }
        // the TERMA requests first, then the density: the two rotations are not live together
        const uint64_t mask = __ballot(need);
        if (mask) {
            if (need) {
                const uint32_t pos = qbase + qdone + (uint32_t)__popcll(mask & lt);
                const v3 gcs = bl_terma_point(a, B, bl, inp);
                TermaReq rq;
                rq.slot = (uint32_t)(meta.off + j);
                rq.beamlet = (uint32_t)(B.beamlet0 + pidx);
                rq.x = gcs.x; rq.y = gcs.y; rq.z = gcs.z;
                if (pos < subcap) qs[pos] = rq;
            }
            qdone += (uint32_t)__popcll(mask);
        }
        if (r + tid < nex) {
            const float h = ok ? 0.5f * (a.delr * bl_density(a, B, bl, inp)) : -1.f;
            Hl[j] = h;
            T[j] = 0.f;
            if (h >= 0.f) {
                nlo = min(nlo, j);
                nhi = max(nhi, j);
            }
        }
    }
    // The reservation counted footprint samples of pass A.2; pass B.2 requests TERMA for those of
    // them it maps to the same pillar again, so any shortfall is closed with empty entries.
    for (uint32_t k = qdone + (uint32_t)tid; k < nfoot; k += 64) {
        const uint32_t pos = qbase + k;
        if (pos < subcap) qs[pos].slot = kNoSlot;
    }
    // Walls before the first and after the last non-wall sample bound no segment and receive no
    // dose: the line's stored interval shrinks to [nlo, nhi], so the walk and the gather skip them
    // (the gather reads samples outside a line's interval as zero, the dose the walk gives a wall).
    for (int o = 32; o > 0; o >>= 1) {
        nlo = min(nlo, __shfl_xor(nlo, o, 64));
        nhi = max(nhi, __shfl_xor(nhi, o, 64));
    }
    if (tid == 0 && (nlo > 0 || nhi < m - 1)) {
        LineMeta t;
        t.off = nhi < nlo ? meta.off : meta.off + (uint32_t)nlo;
        t.i0 = nhi < nlo ? meta.i0 : (int16_t)(i0 + nlo);
        t.len = nhi < nlo ? (int16_t)0 : (int16_t)(nhi - nlo + 1);
        a.lines[line] = t;
    }
    __syncthreads();  // the next line reuses the LDS arrays
    }
    if (a.counts) wave_count(stored, a.counts, blockIdx.x);
}

// Stage 2: the queued samples' TERMA (cudaSiddon_device with the beamlet filter), one sample per
// lane, grid-stride over the queue.
__global__ void __launch_bounds__(256) bl_siddon_kernel(BlArgs a, const TermaReq* __restrict__ q, uint32_t subcap,
                                                        const uint32_t* __restrict__ qcnt, uint32_t* counts) {
    // workgroup b serves sub-queue b % kSubQueues
    const uint32_t sq = blockIdx.x % kSubQueues, per = gridDim.x / kSubQueues;
    const uint32_t n = min(qcnt[sq * kCntStride], subcap);
    const TermaReq* qs = q + (size_t)sq * subcap;
    const v3 ivox = mk3(1.f / a.vd.voxel.x, 1.f / a.vd.voxel.y, 1.f / a.vd.voxel.z);
    uint32_t steps = 0;
    for (uint32_t i = (blockIdx.x / kSubQueues) * 256 + threadIdx.x; i < n; i += per * 256) {
        const TermaReq r = qs[i];
        if (r.slot == kNoSlot) continue;
        const BeamletDev& bl = a.beamlets[r.beamlet];
        const BatchDev& B = a.batches[bl.batch];
        int it = 0;
        a.rev[r.slot] = point_terma_rpl(mk3(r.x, r.y, r.z), B.fmap, B.beam, a.spec, a.vd, ivox, a.dens, a.bhc, bl.fidx,
                                        &it);
        steps += (uint32_t)it;
    }
    if (counts) block_count(steps, counts, blockIdx.x);
}

// A 64-bit value from lane addr / 4 (ds_bpermute on both halves).
__device__ __forceinline__ double bperm_f64(int addr, double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Stage 3, one wave per non-empty REV line: PackRowConvolve (nvbbRayConvolve_device.cu:284-417)
// for every sample.  Both walks of destination j run inside the pillar segment holding j, start at
// j itself (the self sample adds only its first half step) and stop when the radiological
// distance passes the last radius or enters the last radial interval.  Only samples with TERMA
// contribute, and the distance from j to sample k is the difference of the line's radiological
// prefix sums (FP64), so each destination visits just the footprint samples of its segment, in the
// reference's walk order, instead of every sample.
//
// Dynamic LDS: for up to maxn samples {half step, TERMA} (8 B), the radiological centre (FP64),
// the segment bounds and the footprint list (3 x 2 B).
__global__ void __launch_bounds__(64) bl_walk_kernel(BlArgs a, const float* __restrict__ H, int maxn, int len_lo) {
    extern __shared__ double s_dyn[];
    __shared__ float4 sTab[2][kMaxRadii];
    const int tid = threadIdx.x;
    const int nr = a.nradii;
    if (tid < nr) {
        const float* K = a.tables;
        const float* MR = a.tables + (size_t)a.ntheta * nr + 2 * nr;
        for (int d = 0; d < 2; d++) {
            const float* row = K + (size_t)(d == 0 ? a.rowf : a.rowb) * nr;
            const float lo = tid == 0 ? 0.f : MR[tid - 1];
            const float klo = tid == 0 ? 0.f : row[tid - 1];
            sTab[d][tid] = make_float4(lo, klo, (row[tid] - klo) * MR[nr + tid], MR[tid]);
        }
    }
    double* sC = s_dyn;
    float2* sL = reinterpret_cast<float2*>(s_dyn + maxn);
    short* s_lo = reinterpret_cast<short*>(sL + maxn);
    short* s_hi = s_lo + maxn;
    short* s_fp = s_hi + maxn;
    __syncthreads();
    const float rmax = sTab[0][nr - 1].w, rlast = sTab[0][nr - 2].w;
    const float nphi_f = (float)a.nphi;
    uint32_t iters = 0;
    const uint64_t G = gridDim.x;
    for (uint64_t base = xcd_remap(blockIdx.x, gridDim.x); base < a.nlines; base += 64 * G) {
    LineMeta mine{};
    if (base + (uint64_t)tid * G < a.nlines) mine = a.lines[base + (uint64_t)tid * G];
    // empty lines and other launches' lines are skipped here
    uint64_t todo = __ballot(mine.len >= len_lo && mine.len <= maxn);
    while (todo) {
        const int kl = __ffsll((unsigned long long)todo) - 1;
        todo &= todo - 1ull;
        LineMeta meta = line_of_lane(mine, kl);
        meta.off = __builtin_amdgcn_readfirstlane(meta.off);  // uniform: scalar registers
        const int m = __builtin_amdgcn_readfirstlane((int)meta.len);
        // the scans' lane arithmetic is redone per line from an opaque lane id: hoisted out of
        // the line loop, it held some 20 VGPRs for the whole kernel
        int ln = tid;
        __asm__ volatile("" : "+v"(ln));
        float* T = a.rev + meta.off;
        const float* Hl = H + meta.off;
        for (int j = tid; j < m; j += 64) sL[j] = make_float2(Hl[j], T[j]);
        __syncthreads();
        // chunked scans: radiological prefix, nearest walls, footprint list
        const int per = (m + 63) / 64;
        const int ca = min(m, tid * per), cb = min(m, ca + per);
        double csum = 0.0;
        int last_wall = -1, first_wall = m, nfp = 0;
        for (int k = ca; k < cb; k++) {
            const float2 e = sL[k];
            if (e.x < 0.f) last_wall = k;
            else csum += 2.0 * (double)e.x;
            if (e.y != 0.f) nfp++;
        }
        for (int k = cb - 1; k >= ca; k--)
            if (sL[k].x < 0.f) first_wall = k;
        double pc = csum;
        int pre = last_wall, suf = first_wall, pf = nfp;
        for (int o = 1; o < 64; o <<= 1) {
            const int iu = max(ln - o, 0) << 2, id = min(ln + o, 63) << 2;
            const double uc = bperm_f64(iu, pc);
            const int up = __builtin_amdgcn_ds_bpermute(iu, pre);
            const int uf = __builtin_amdgcn_ds_bpermute(iu, pf);
            const int dn = __builtin_amdgcn_ds_bpermute(id, suf);
            if (tid >= o) { pc += uc; pre = max(pre, up); pf += uf; }
            if (tid + o < 64) suf = min(suf, dn);
        }
        double cacc = pc - csum;
        int left = __builtin_amdgcn_ds_bpermute(max(ln - 1, 0) << 2, pre);
        int right = __builtin_amdgcn_ds_bpermute(min(ln + 1, 63) << 2, suf);
        int fpos = pf - nfp;
        const int nfoot = __builtin_amdgcn_readlane(pf, 63);
        if (tid == 0) left = -1;
        if (tid == 63) right = m;
        for (int k = ca; k < cb; k++) {
            const float2 e = sL[k];
            const double hh = e.x < 0.f ? 0.0 : (double)e.x;
            sC[k] = cacc + hh;
            cacc += 2.0 * hh;
            if (e.x < 0.f) left = k;
            s_lo[k] = (short)(left + 1);
            if (e.y != 0.f) s_fp[fpos++] = (short)k;
        }
        for (int k = cb - 1; k >= ca; k--) {
            if (sL[k].x < 0.f) right = k;
            s_hi[k] = (short)(right - 1);
        }
        __syncthreads();
        for (int base = 0; base < m; base += 64) {
            const int j = base + tid;
            if (j >= m) break;
            const float2 ej = sL[j];
            float sum = 0.f;
            if (ej.x >= 0.f && nfoot > 0) {
                const int slo = s_lo[j], shi = s_hi[j];
                const double cj = sC[j];
                int lo = 0, hi = nfoot;  // first footprint entry >= j
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (s_fp[mid] < j) lo = mid + 1; else hi = mid;
                }
                const int p = lo;
                const float r0 = ej.x;
                for (int d = 0; d < 2; d++) {
                    const float4* tab = sTab[d];
                    int s = 0;
                    float4 iv = tab[0];
                    if (!(r0 <= rmax)) continue;
                    sum += kc_at(r0, tab, nr, s, iv) * ej.y;  // the self sample
                    iters++;
                    if (r0 > rlast) continue;
                    if (d == 0) {
                        for (int q = p - 1; q >= 0; q--) {
                            const int k = s_fp[q];
                            if (k < slo) break;
                            const double ck = sC[k];
                            const float rp = k + 1 == j ? r0 : (float)(cj - sC[k + 1]);
                            if (rp > rlast) break;
                            const float rk = (float)(cj - ck);
                            if (rk > rmax) break;
                            const float kp = kc_at(rp, tab, nr, s, iv);
                            const float kk = kc_at(rk, tab, nr, s, iv);
                            sum += (kk - kp) * sL[k].y;
                            iters++;
                        }
                    } else {
                        for (int q = (p < nfoot && s_fp[p] == j) ? p + 1 : p; q < nfoot; q++) {
                            const int k = s_fp[q];
                            if (k > shi) break;
                            const double ck = sC[k];
                            const float rp = k - 1 == j ? r0 : (float)(sC[k - 1] - cj);
                            if (rp > rlast) break;
                            const float rk = (float)(ck - cj);
                            if (rk > rmax) break;
                            const float kp = kc_at(rp, tab, nr, s, iv);
                            const float kk = kc_at(rk, tab, nr, s, iv);
                            sum += (kk - kp) * sL[k].y;
                            iters++;
                        }
                    }
                }
            }
            T[j] = ej.x >= 0.f ? sum / nphi_f : 0.f;
        }
        __syncthreads();  // the next line reuses the LDS arrays
    }
    }
    if (a.counts) wave_count(iters, a.counts, blockIdx.x);
}

// Only REV lines with footprint samples in a pillar carry dose in it, and those lines pass within
// sqrt(2)*gw of that pillar's beamlet axis (footprint samples satisfy |x-fc|,|z-fc| <= gw, the
// farthest reach of the footprint bounds, itself <= fw).  The four lines a voxel reads lie within
// sqrt(2) of the line through it, so if that line passes farther than sqrt(2)*(gw + 1) + 0.5 from
// the axes of the voxel's pillar and of the neighbouring pillars it touches, the fetch is zero and
// is skipped.  Depends on pX and pZ only.
__device__ __forceinline__ bool bl_gather_hit(const BatchDev& B, const BlConeDev& c, int pX, int pZ) {
    if (!(B.fw < 1e29f)) return true;
    const float wx = c.kinv_m.m[1], wz = c.kinv_m.m[7];
    const float wn = sqrtf(wx * wx + wz * wz);
    const float bound = 1.41421356f * (B.gw + 1.f) + 0.5f;
    const int pix = pX / B.pd[0], piy = pZ / B.pd[2];
    const int rx = pX - pix * B.pd[0], rz = pZ - piy * B.pd[2];
    const int sx = rx < 2 ? -1 : (rx >= B.pd[0] - 2 ? 1 : 0);
    const int sz = rz < 2 ? -1 : (rz >= B.pd[2] - 2 ? 1 : 0);
    bool hit = false;
    for (int ox = 0; ox <= (sx != 0 ? 1 : 0) && !hit; ox++)
        for (int oz = 0; oz <= (sz != 0 ? 1 : 0) && !hit; oz++) {
            const int qx = pix + ox * sx, qz = piy + oz * sz;
            if (qx < 0 || qz < 0 || qx >= B.npx || qz >= B.npy || qx + qz * B.npx >= B.n) continue;
            const float dx = (float)(pX - qx * B.pd[0]) - B.fc[0], dz = (float)(pZ - qz * B.pd[2]) - B.fc[1];
            const float dist = wn > 1e-6f ? fabsf(dx * wz - dz * wx) / wn : sqrtf(dx * dx + dz * dz);
            hit = dist <= bound;
        }
    return hit;
}

// PackedREVtoBEVdose (nvbbRayConvolve_device.cu:427-452) for one cone direction over every batch's
// packed grid: dose += cone dose where it is positive.  One wave per 16x4x4 voxel tile (pY
// fastest: compact in every direction, so neighbouring lanes read neighbouring REV lines); each
// lane owns one (pY, pZ) and the tile's four pX voxels.  The four voxels' loads are all issued
// before any of their stores, so a lane has four fetches in flight, and the tile's scalar set-up
// is paid once per wave.
constexpr int kGatherX = 4;  // pX voxels per lane (the tile's depth along pX)
__global__ void __launch_bounds__(64) bl_gather_kernel(BlArgs a, uint32_t ntiles) {
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const int lo = a.tile_batch ? (int)a.tile_batch[tile] : 0;
    const BatchDev& B = a.batches[lo];
    const BlConeDev& c = a.cones[lo];
    const uint32_t t = tile - B.tile0;
    const int ty = (int)(t % (uint32_t)B.tiles[0]);
    const int tz = (int)((t / (uint32_t)B.tiles[0]) % (uint32_t)B.tiles[1]);
    const int tx = (int)(t / ((uint32_t)B.tiles[0] * (uint32_t)B.tiles[1]));
    const int tid = threadIdx.x;
    const int pY = ty * 16 + (tid & 15), pZ = tz * 4 + (tid >> 4);
    if (pY >= B.gd[1] || pZ >= B.gd[2]) return;
    float old[kGatherX], dc[kGatherX];
    float* dst[kGatherX];
#pragma unroll
    for (int s = 0; s < kGatherX; s++) {
        const int pX = tx * kGatherX + s;
        dc[s] = 0.f;
        old[s] = 0.f;
        dst[s] = nullptr;
        if (pX >= B.gd[0] || !bl_gather_hit(B, c, pX, pZ)) continue;
        // continuous resampling: the matrix form of rotateKernel (the Rodrigues form differs only
        // in the last bits of the fetch coordinate)
        const v3 rk = c.kfwd_m(mk3((float)pX, (float)pY, (float)pZ));
        const v3 ri = mk3(fmaf(-c.minc.x, a.rrv.x, rk.x), fmaf(-c.minc.y, a.rrv.y, rk.y),
                          fmaf(-c.minc.z, a.rrv.z, rk.z));
        // the voxel's packed dose is read before the fetch so that its latency overlaps the
        // fetch's (this lane alone writes the voxel in this launch)
        dst[s] = a.pdose + B.pdose + (size_t)pY + (size_t)B.gd[1] * ((size_t)pZ + (size_t)B.gd[2] * pX);
        old[s] = *dst[s];
        dc[s] = rev_fetch(a.lines, a.rev, c.line0, c.nx, c.ny, c.nz, ri.y + 0.5f, ri.z + 0.5f, ri.x + 0.5f);
    }
#pragma unroll
    for (int s = 0; s < kGatherX; s++)
        if (dc[s] > 0.f) *dst[s] = old[s] + dc[s];
}

// The batch of every gather tile of a job (one fill per batch range).
__global__ void __launch_bounds__(256) bl_tile_batch_kernel(const BatchDev* __restrict__ batches,
                                                            uint32_t* __restrict__ tb) {
    const int b = blockIdx.y;
    const BatchDev& B = batches[b];
    const uint32_t n = (uint32_t)B.tiles[0] * (uint32_t)B.tiles[1] * (uint32_t)B.tiles[2];
    for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256) tb[B.tile0 + t] = (uint32_t)b;
}

__global__ void __launch_bounds__(256) bl_fill_kernel(float* __restrict__ p, uint64_t n, float v) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) p[i] = v;
}

struct UnpackArgs {
    const BatchDev* batches;
    const BeamletDev* beamlets;
    const float* pdose;
    float* arena;
    int nbeamlets;
    uint64_t narena;
    v3 start, voxel, rvs;
    int bbs[3];   // calc_bbox start
};

// UnpackBEVDosePillar (nvbbRayConvolve_device.cu:459-521) over each beamlet's voxel box: the
// calc_bbox voxel goes to its pillar (beam then beamlet rotation), rim and walls read nothing, and
// the packed dose is fetched trilinearly.  Arena order inside a box: i outer, j, k inner.
__global__ void __launch_bounds__(256) bl_unpack_kernel(UnpackArgs a) {
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= a.narena) return;
    int lo = 0, hi = a.nbeamlets - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.beamlets[mid].arena <= e) lo = mid; else hi = mid - 1;
    }
    const BeamletDev& bl = a.beamlets[lo];
    const BatchDev& B = a.batches[bl.batch];
    const uint64_t le = e - bl.arena;
    const int k = (int)(le % (uint64_t)bl.boxn[2]);
    const int j = (int)((le / (uint64_t)bl.boxn[2]) % (uint64_t)bl.boxn[1]);
    const int i = (int)(le / ((uint64_t)bl.boxn[2] * bl.boxn[1]));
    const int tX = bl.box0[0] + i, tY = bl.box0[1] + j, tZ = bl.box0[2] + k;
    const v3 gc = mk3(((float)tX + (float)a.bbs[0]) * a.voxel.x + a.start.x,
                      ((float)tY + (float)a.bbs[1]) * a.voxel.y + a.start.y,
                      ((float)tZ + (float)a.bbs[2]) * a.voxel.z + a.start.z);
    v3 pidx = rot1(bl.fwd, rot2(B.beam.beam_fwd, gc - bl.pst)) / a.rvs + 0.5f;
    float v = 0.f;
    const float lo_f = (float)(kWall + kBuffer);
    if (!(pidx.x < lo_f || pidx.y < lo_f || pidx.z < lo_f || pidx.x >= (float)(B.pd[0] - kBuffer) ||
          pidx.y >= (float)(B.pd[1] - kBuffer) || pidx.z >= (float)(B.pd[2] - kBuffer))) {
        pidx = mk3(pidx.x + (float)(bl.pix * B.pd[0]), pidx.y, pidx.z + (float)(bl.piy * B.pd[2]));
        const float dc = tex3_border(a.pdose + B.pdose, B.gd[1], B.gd[2], B.gd[0], pidx.y, pidx.z, pidx.x);
        if (dc > 0.f) v = dc;
    }
    a.arena[e] = v;
}

// ---- sparse columns without an arena -------------------------------------------------------------
// One workgroup per beamlet walks the calc_bbox rows (i, j) of its box in the emission order of
// _sparsify (i outer, j, k inner).  The pillar's voxels along a row form one interval of k (the
// pillar is a box in its own frame and the frame is affine in k), found by clipping the row
// against the pillar's three slabs (matrix form, widened); inside it the exact unpack test and
// fetch of UnpackBEVDosePillar decide each value.  Nothing is stored between passes: the values
// are recomputed for the max, the counts and the emission.

struct ColArgs {
    const BatchDev* batches;
    const BeamletDev* beamlets;
    const float* pdose;
    v3 start, voxel, rvs;
    int bbs[3];        // calc_bbox start
    int nx, ny;        // full volume dims (keys)
    float thresh;
    uint32_t* stat;    // per beamlet {kept, dropped, max bits, 0}
    uint16_t* rowcnt;  // kept entries of every (i, j) row of every beamlet's box (from the scan)
    const uint64_t* off;
    uint64_t* keys;
    float* vals;
};

// UnpackBEVDosePillar (nvbbRayConvolve_device.cu:459-521) at one calc_bbox voxel.
__device__ __forceinline__ float pillar_value(const ColArgs& a, const BeamletDev& bl, const BatchDev& B, int tX, int tY,
                                              int tZ) {
    const v3 gc = mk3(((float)tX + (float)a.bbs[0]) * a.voxel.x + a.start.x,
                      ((float)tY + (float)a.bbs[1]) * a.voxel.y + a.start.y,
                      ((float)tZ + (float)a.bbs[2]) * a.voxel.z + a.start.z);
    v3 pidx = rot1(bl.fwd, rot2(B.beam.beam_fwd, gc - bl.pst)) / a.rvs + 0.5f;
    const float lo_f = (float)(kWall + kBuffer);
    if (pidx.x < lo_f || pidx.y < lo_f || pidx.z < lo_f || pidx.x >= (float)(B.pd[0] - kBuffer) ||
        pidx.y >= (float)(B.pd[1] - kBuffer) || pidx.z >= (float)(B.pd[2] - kBuffer))
        return 0.f;
    pidx = mk3(pidx.x + (float)(bl.pix * B.pd[0]), pidx.y, pidx.z + (float)(bl.piy * B.pd[2]));
    const float dc = tex3_border(a.pdose + B.pdose, B.gd[1], B.gd[2], B.gd[0], pidx.y, pidx.z, pidx.x);
    return dc > 0.f ? dc : 0.f;
}

// k range [k0, k1) of row (tX, tY) that can lie inside the pillar's unpacked region.
__device__ __forceinline__ void pillar_row(const ColArgs& a, const BeamletDev& bl, const BatchDev& B, int tX, int tY,
                                           int& k0, int& k1) {
    const v3 g0 = mk3(((float)tX + (float)a.bbs[0]) * a.voxel.x + a.start.x,
                      ((float)tY + (float)a.bbs[1]) * a.voxel.y + a.start.y,
                      ((float)bl.box0[2] + (float)a.bbs[2]) * a.voxel.z + a.start.z);
    const v3 q0 = bl.fwd_m(g0 - bl.pst) / a.rvs + 0.5f;
    const v3 dq = mk3(bl.fwd_m.m[2] * a.voxel.z / a.rvs.x, bl.fwd_m.m[5] * a.voxel.z / a.rvs.y,
                      bl.fwd_m.m[8] * a.voxel.z / a.rvs.z);
    float t0 = 0.f, t1 = (float)bl.boxn[2];
    const float lo_f = (float)(kWall + kBuffer);
    clip_slab(q0.x, dq.x, lo_f, (float)(B.pd[0] - kBuffer), t0, t1);
    clip_slab(q0.y, dq.y, lo_f, (float)(B.pd[1] - kBuffer), t0, t1);
    clip_slab(q0.z, dq.z, lo_f, (float)(B.pd[2] - kBuffer), t0, t1);
    k0 = 0;
    k1 = 0;
    if (t0 <= t1) {
        k0 = max(0, (int)floorf(t0) - 1);
        k1 = min(bl.boxn[2], (int)ceilf(t1) + 2);
        if (k1 < k0) k1 = k0;
    }
}

// Per-beamlet column statistics of _sparsify (DoseCalcIO/sparse_io.cpp:22-58): max over the
// column (zeros outside the pillar count), kept (v > t * max) and dropped (0 < v <= t * max).
__global__ void __launch_bounds__(256) bl_colscan_kernel(ColArgs a) {
    __shared__ float s_max[4];
    __shared__ uint32_t s_k[4], s_d[4];
    const BeamletDev& bl = a.beamlets[blockIdx.x];
    const BatchDev& B = a.batches[bl.batch];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int rows = bl.boxn[0] * bl.boxn[1];
    float mx = 0.f;
    for (int r = t; r < rows; r += 256) {
        const int i = r / bl.boxn[1], j = r - i * bl.boxn[1];
        int k0, k1;
        pillar_row(a, bl, B, bl.box0[0] + i, bl.box0[1] + j, k0, k1);
        for (int k = k0; k < k1; k++)
            mx = fmaxf(mx, pillar_value(a, bl, B, bl.box0[0] + i, bl.box0[1] + j, bl.box0[2] + k));
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) s_max[w] = mx;
    __syncthreads();
    const float cmax = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
    const float cut = a.thresh * cmax;
    uint32_t kept = 0, dropped = 0;
    for (int r = t; r < rows; r += 256) {
        const int i = r / bl.boxn[1], j = r - i * bl.boxn[1];
        int k0, k1;
        pillar_row(a, bl, B, bl.box0[0] + i, bl.box0[1] + j, k0, k1);
        uint32_t rk = 0;
        for (int k = k0; k < k1; k++) {
            const float v = pillar_value(a, bl, B, bl.box0[0] + i, bl.box0[1] + j, bl.box0[2] + k);
            if (v > cut) rk++;
            else if (v > 0.f) dropped++;
        }
        kept += rk;
        a.rowcnt[(size_t)bl.row0 + (size_t)r] = (uint16_t)rk;
    }
    for (int o = 32; o > 0; o >>= 1) {
        kept += __shfl_xor(kept, o, 64);
        dropped += __shfl_xor(dropped, o, 64);
    }
    if (lane == 0) { s_k[w] = kept; s_d[w] = dropped; }
    __syncthreads();
    if (t == 0) {
        a.stat[4 * blockIdx.x + 0] = s_k[0] + s_k[1] + s_k[2] + s_k[3];
        a.stat[4 * blockIdx.x + 1] = s_d[0] + s_d[1] + s_d[2] + s_d[3];
        a.stat[4 * blockIdx.x + 2] = __float_as_uint(cmax);
        a.stat[4 * blockIdx.x + 3] = 0;
    }
}

// Exclusive scan of the kept counts over the job's beamlets (one workgroup; n is small).
__global__ void __launch_bounds__(1024) bl_offsets_kernel(uint32_t* __restrict__ stat, int n, uint64_t* __restrict__ off) {
    __shared__ uint64_t s[1024];
    uint64_t carry = 0;
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        const uint64_t v = i < n ? stat[4 * i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const uint64_t add = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < n) off[i] = carry + s[threadIdx.x] - v;
        carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) off[n] = carry;
}

// Writes each beamlet's kept entries at its offset (one workgroup per beamlet), key = full-volume
// linear index, value = dose, in the emission order of _sparsify.  The rows go in chunks of 256,
// one per thread, with their kept counts from the column scan; each wave's 64 rows form one
// contiguous range of the output.  A wave fills that range window by window: every lane evaluates
// its row's values once, in k order, and stages the kept entries that fall in the window in LDS;
// then the wave copies the window out with whole-wave, contiguous stores.
constexpr uint32_t kEmitWindow = 512;  // staged entries per wave and round (6 KB)
__global__ void __launch_bounds__(256) bl_emit_kernel(ColArgs a) {
    __shared__ uint64_t s_key[4][kEmitWindow];
    __shared__ float s_val[4][kEmitWindow];
    __shared__ uint32_t s_w[4];
    const BeamletDev& bl = a.beamlets[blockIdx.x];
    const BatchDev& B = a.batches[bl.batch];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int rows = bl.boxn[0] * bl.boxn[1];
    const float cut = a.thresh * __uint_as_float(a.stat[4 * blockIdx.x + 2]);
    uint64_t pos = a.off[blockIdx.x];
    for (int base = 0; base < rows; base += 256) {
        const int r = base + t;
        const uint32_t n = r < rows ? (uint32_t)a.rowcnt[(size_t)bl.row0 + (size_t)r] : 0u;
        // inclusive scan of n over the wave, and the waves' totals
        uint32_t inc = n;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(inc, o, 64);
            if (lane >= o) inc += u;
        }
        if (lane == 63) s_w[w] = inc;
        __syncthreads();
        uint32_t before = 0, total = 0, most = 0;
        for (int q = 0; q < 4; q++) {
            before += q < w ? s_w[q] : 0u;
            total += s_w[q];
            most = max(most, s_w[q]);
        }
        const uint32_t W = s_w[w];
        __syncthreads();
        int X = 0, Y = 0, k = 0, k1 = 0;
        if (n) {
            const int i = r / bl.boxn[1], j = r - i * bl.boxn[1];
            pillar_row(a, bl, B, bl.box0[0] + i, bl.box0[1] + j, k, k1);
            X = bl.box0[0] + i;
            Y = bl.box0[1] + j;
        }
        uint32_t p = inc - n;  // next output of this lane, relative to the wave's range
        const uint64_t wout = pos + before;
        // the same number of rounds in every wave (the barriers), enough for the fullest wave
        for (uint32_t wb = 0; wb < most; wb += kEmitWindow) {
            const uint32_t we = min(W, wb + kEmitWindow);
            while (p < inc && p < we && k < k1) {
                const float v = pillar_value(a, bl, B, X, Y, bl.box0[2] + k);
                if (v > cut) {
                    s_key[w][p - wb] = ((uint64_t)(bl.box0[2] + k + a.bbs[2]) * (uint64_t)a.ny +
                                        (uint64_t)(Y + a.bbs[1])) * (uint64_t)a.nx + (uint64_t)(X + a.bbs[0]);
                    s_val[w][p - wb] = v;
                    p++;
                }
                k++;
            }
            __syncthreads();
            for (uint32_t e = wb + (uint32_t)lane; e < we; e += 64) {
                a.keys[wout + e] = s_key[w][e - wb];
                a.vals[wout + e] = s_val[w][e - wb];
            }
            __syncthreads();
        }
        pos += total;
    }
}

// ------------------------------------------------------------------------------------------------
// plain path: the reference dataflow on HIP, kept as the baseline the batched path is measured
// against (dosecalc-beamlet/nvbbRayConvolve_host.cu:692-1409): one batch at a time; per cone a
// raytrace over the whole REV grid, PackRowConvolve with one thread per sample, and the
// packed-BEV gather; then per beamlet a full calc_bbox unpack, a device-to-host copy of the
// dense column and a host sparsify.

// cudaBeamletRaytrace (nvbbRayConvolve_device.cu:19-242), one thread per REV sample: density
// (-1e-12 for walls and outside) and the Siddon TERMA of the sample's beamlet.
__global__ void __launch_bounds__(256) plain_ray_kernel(BlArgs a, float* __restrict__ dens, float* __restrict__ terma) {
    const BlConeDev& c = a.cones[0];
    const BatchDev& B = a.batches[c.batch];
    const uint64_t n = (uint64_t)c.nx * c.ny * c.nz;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int tX = (int)(i % (uint64_t)c.nx);
    const uint64_t l = i / (uint64_t)c.nx;
    const int tY = (int)(l % (uint64_t)c.ny), tZ = (int)(l / (uint64_t)c.ny);
    float d = -1e-12f, t = 0.f;
    int pidx;
    v3 inp;
    const v3 ipd = mk3(1.f / (float)B.pd[0], 1.f / (float)B.pd[1], 1.f / (float)B.pd[2]);
    if (bl_locate(a, c, B, ipd, tX, tY, tZ, pidx, inp)) {
        const PillarRec bl = pillar_rec(a.beamlets[B.beamlet0 + pidx]);
        t = point_terma(bl_terma_point(a, B, bl, inp), B.fmap, B.beam, a.spec, a.vd, a.dens, a.bhc, bl.fidx, nullptr);
        d = bl_density(a, B, bl, inp);
    }
    dens[i] = d;
    terma[i] = t;
}

// PackRowConvolve body for one destination (nvbbRayConvolve_device.cu:284-417): two walks that
// stop at a wall, past the last radius, or after entering the last radial interval.
__device__ float plain_walk(const float* __restrict__ d, const float* __restrict__ T, int n, int tx,
                            const float* __restrict__ R, const float* __restrict__ IW, const float* __restrict__ rowF,
                            const float* __restrict__ rowB, int nr, float delr) {
    if (d[tx] < 0.f) return 0.f;
    float sum = 0.f;
    for (int dir = -1; dir <= 1; dir += 2) {
        const float* row = dir < 0 ? rowF : rowB;
        int k = tx, s = 0;
        float r = 0.f, last = 0.f;
        do {
            if (k < 0 || k >= n || d[k] < 0.f) break;
            const float h = 0.5f * (delr * d[k]);
            r += h;
            if (r > R[nr - 1]) break;
            while (R[s] < r) s++;
            const float lo = s == 0 ? 0.f : R[s - 1];
            const float klo = s == 0 ? 0.f : row[s - 1];
            const float cum = klo + (r - lo) * ((row[s] - klo) * IW[s]);
            sum += (cum - last) * T[k];
            last = cum;
            if (k != tx) r += h;
            k += dir;
        } while (s < nr - 1);
    }
    return sum;
}

// One workgroup per REV line, the line's density and TERMA in LDS.
__global__ void __launch_bounds__(256) plain_conv_kernel(BlArgs a, const float* __restrict__ dens,
                                                         const float* __restrict__ terma, float* __restrict__ out) {
    extern __shared__ float s_pl[];
    __shared__ float sR[kMaxRadii], sIW[kMaxRadii], sF[kMaxRadii], sB[kMaxRadii];
    const BlConeDev& c = a.cones[0];
    const int n = c.nx, tid = threadIdx.x, nr = a.nradii;
    const uint64_t base = (uint64_t)blockIdx.x * n;
    for (int j = tid; j < n; j += 256) {
        s_pl[j] = dens[base + j];
        s_pl[n + j] = terma[base + j];
    }
    if (tid < nr) {
        const float* MR = a.tables + (size_t)a.ntheta * nr + 2 * nr;
        sR[tid] = MR[tid];
        sIW[tid] = MR[nr + tid];
        sF[tid] = a.tables[(size_t)a.rowf * nr + tid];
        sB[tid] = a.tables[(size_t)a.rowb * nr + tid];
    }
    __syncthreads();
    for (int j = tid; j < n; j += 256)
        out[base + j] = plain_walk(s_pl, s_pl + n, n, j, sR, sIW, sF, sB, nr, a.delr) / (float)a.nphi;
}

// Line table of a dense REV grid (every sample stored), so the packed-BEV gather reads it too.
__global__ void __launch_bounds__(256) plain_meta_kernel(LineMeta* __restrict__ lines, uint64_t nlines, int nx) {
    const uint64_t l = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (l >= nlines) return;
    LineMeta m;
    m.i0 = 0;
    m.len = (int16_t)nx;
    m.off = (uint32_t)(l * (uint64_t)nx);
    lines[l] = m;
}

// ------------------------------------------------------------------------------------------------
// host planning

namespace {

struct Rev {
    v3 minc;
    uint32_t size[3];
};

// findREV (dosecalc-beamlet/nvbbRayConvolve_host.cu:104-173): REV grid holding the rotated box
// [0, lim]; false when it exceeds max_rev (the reference's ArraySizeError).
bool find_rev(const cccs_problem& p, const uint32_t* max_rev, v3 lim, float theta, float phi, Rev& out) {
    Rot2 kf = rot2_kernel_fwd(theta, phi);
    v3 mn = mk3(FLT_MAX, FLT_MAX, FLT_MAX), mx = mk3(FLT_MIN, FLT_MIN, FLT_MIN);
    for (int c = 0; c < 8; c++) {
        v3 in = mk3(lim.x * (float)(c >> 2 & 1), lim.y * (float)(c >> 1 & 1), lim.z * (float)(c & 1));
        v3 o = rot2(kf, in);
        mn.x = std::min(mn.x, o.x); mn.y = std::min(mn.y, o.y); mn.z = std::min(mn.z, o.z);
        mx.x = std::max(mx.x, o.x); mx.y = std::max(mx.y, o.y); mx.z = std::max(mx.z, o.z);
    }
    out.size[0] = (uint32_t)std::ceil(std::fabs(mx.y - mn.y) / p.rev_longspacing);
    out.size[1] = (uint32_t)std::ceil(std::fabs(mx.z - mn.z) / p.rev_latspacing);
    out.size[2] = (uint32_t)std::ceil(std::fabs(mx.x - mn.x) / p.rev_latspacing);
    out.minc = mn;
    return !(out.size[0] > max_rev[0] || out.size[1] > max_rev[1] || out.size[2] > max_rev[2]);
}

float comp(v3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// One active beamlet of a beam.
struct BeamletPlan {
    int fidx;
    v3 start, end;
    float theta, phi;  // divergence angles of the beamlet axis
    float len;
    v3 pst;
};

struct BeamPlan {
    int beam;                     // position in the request
    const cccs_beam* b;
    std::vector<BeamletPlan> bl;
    int pd[3];
    v3 psize;                     // pillar size [cm]
    float max_bs;                 // largest divergent beamlet width at a beamlet's exit [cm]
    float min_exit;               // smallest source-to-exit distance [cm]
    std::vector<int> bsizes;      // beamlets per batch
};

// calcBeamletAnchors (dosecalc-beamlet/nvbbRayConvolve_host.cu:180-280): beamlet isocentre,
// divergence angles and the entry/exit points of the beamlet axis through the calc_bbox.
bool beamlet_anchor(const cccs_problem& p, const cccs_beam& b, int idx, BeamletPlan& o, v3& biso) {
    const v3 src = ld3(b.source), iso = ld3(b.isocenter);
    const int fx = (int)b.fmap_size[0], fy = (int)b.fmap_size[1];
    const int bx = idx % fx, by = idx / fx;
    const v3 bdiff = mk3(b.beamlet_size[0] * (-0.5f * (float)fx + (float)bx + 0.5f), 0.f,
                         b.beamlet_size[1] * (-0.5f * (float)fy + (float)by + 0.5f));
    biso = iso + rot2(rot2_beam_inv(b.azimuth, b.zenith, b.coll), bdiff);
    o.theta = std::acos(len3(iso - src) / len3(biso - src));
    o.phi = std::atan2(bdiff.z, bdiff.x);
    const v3 voxel = ld3(p.voxel), dstart = ld3(p.start);
    const v3 bb_begin = dstart + voxel * ldu3(p.bbox_start);
    const v3 bb_end = bb_begin + voxel * ldu3(p.bbox_size);
    const double buffer = 0.05;
    const v3 d = biso - src;
    v3 hits[6];
    for (int ax = 0; ax < 3; ax++) {
        const double den = (double)comp(d, ax) + 1e-12;
        const double a1 = (double)(comp(bb_begin, ax) - comp(src, ax)) / den;
        const double a2 = (double)(comp(bb_end, ax) - comp(src, ax)) / den;
        hits[2 * ax] = src + (float)a1 * d;
        hits[2 * ax + 1] = src + (float)a2 * d;
    }
    double dmin = std::numeric_limits<double>::max(), dmax = std::numeric_limits<double>::min();
    o.start = mk3(FLT_MAX, FLT_MAX, FLT_MAX);
    o.end = mk3(FLT_MIN, FLT_MIN, FLT_MIN);
    for (int ax = 0; ax < 3; ax++)
        for (int s = 0; s < 2; s++) {
            const v3 h = hits[2 * ax + s];
            const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
            if (comp(h, a1) + buffer >= comp(bb_begin, a1) && comp(h, a1) - buffer <= comp(bb_end, a1) &&
                comp(h, a2) + buffer >= comp(bb_begin, a2) && comp(h, a2) - buffer <= comp(bb_end, a2)) {
                const float dist = len3(h - src);
                if (dist < dmin) { o.start = h; dmin = len3(o.start - src); }
                if (dist > dmax) { o.end = h; dmax = len3(o.end - src); }
            }
        }
    return dmin < (double)std::numeric_limits<float>::max() && dmax > (double)std::numeric_limits<float>::min();
}

// Pillar sizes, origins and batch sizes of one beam (host.cu:575-834).
int plan_beam(const cccs_ctx* ctx, const cccs_beam& b, const float* fmap, int beam_index, BeamPlan& P) {
    const cccs_problem& p = ctx->prob;
    P.beam = beam_index;
    P.b = &b;
    P.bl.clear();
    P.bsizes.clear();
    const int nf = (int)(b.fmap_size[0] * b.fmap_size[1]);
    const v3 rvs = mk3(p.rev_latspacing, p.rev_longspacing, p.rev_latspacing);
    const v3 src = ld3(b.source);
    float max_len = 0.f, mbx = 0.f, mby = 0.f, min_exit = FLT_MAX;
    for (int i = 0; i < nf; i++) {
        if (!(fmap[i] > 0.f)) continue;
        BeamletPlan o;
        o.fidx = i;
        v3 biso;
        if (!beamlet_anchor(p, b, i, o, biso)) return set_error(CCCS_EGEOM, "beamlet axis misses the calc_bbox");
        o.len = len3(o.end - o.start);
        max_len = std::max(max_len, o.len);
        const float num = len3(o.end - src), den = len3(biso - src);
        mbx = std::max(mbx, b.beamlet_size[0] * num / den);
        mby = std::max(mby, b.beamlet_size[1] * num / den);
        min_exit = std::min(min_exit, num);
        P.bl.push_back(o);
    }
    const int n = (int)P.bl.size();
    if (n == 0) return CCCS_OK;
    const float ke = p.kernel_extent;
    const float plong = max_len + 2.f * ke + (float)kWall * rvs.y;
    const float ptx = mbx + 2.f * ke + (float)kWall * rvs.x;
    const float pty = mby + 2.f * ke + (float)kWall * rvs.x;
    const v3 expand = rvs - mk3(std::fmod(ptx, rvs.x), std::fmod(plong, rvs.y), std::fmod(pty, rvs.z));
    const v3 psize = mk3(ptx, plong, pty) + expand + 2.f * (float)kBuffer * rvs;
    const v3 pdf = psize / rvs;
    P.pd[0] = (int)pdf.x; P.pd[1] = (int)pdf.y; P.pd[2] = (int)pdf.z;
    P.psize = psize;
    P.max_bs = std::max(mbx, mby);
    P.min_exit = min_exit;
    const Rot2 binv = rot2_beam_inv(b.azimuth, b.zenith, b.coll);
    for (BeamletPlan& o : P.bl) {
        v3 go = mk3(-0.5f * (psize.x + (float)kWall * rvs.x), -0.5f * (psize.y - o.len + (float)kWall * rvs.y),
                    -0.5f * (psize.z + (float)kWall * rvs.z));
        go = rot2(binv, rot1(beamlet_rot(o.theta, o.phi, true), go));
        o.pst = o.start + go;
    }
    uint32_t mr[3] = {p.max_rev_size[0], p.max_rev_size[1], p.max_rev_size[2]};
    if (mr[0] == 0) mr[0] = mr[1] = mr[2] = 500;  // dosecalc-preprocess/argparser.h:29
    for (int nb = 1;; nb++) {
        if (nb > kMaxBatches) return set_error(CCCS_ENOMEM, "beamlets do not fit max_rev_size in 100 batches");
        P.bsizes.assign(nb, n / nb);
        for (int r = 0, rem = n - (n / nb) * nb; rem > 0; r++, rem--) P.bsizes[r % nb]++;
        bool ok = true;
        for (int bi = 0; bi < nb && ok; bi++) {
            const int m = P.bsizes[bi];
            const int npx = (int)std::ceil(std::sqrt((float)m));
            const int npy = (int)std::ceil((float)m / (float)npx);
            const v3 gsize = mk3((float)(npx * P.pd[0] + kWall), (float)(P.pd[1] + kWall), (float)(npy * P.pd[2] + kWall)) * rvs;
            for (int r = 0; r < ctx->nrays && ok; r++) {
                Rev rv;
                const float th = ctx->theta[r] * 3.14159265358979323846f / 180.f;
                const float ph = ctx->phi[r] * 3.14159265358979323846f / 180.f;
                if (!find_rev(p, mr, gsize, th, ph, rv)) ok = false;
                else if (rv.size[0] > (uint32_t)kMaxLine) return set_error(CCCS_EGEOM, "REV line longer than 1024 samples");
            }
        }
        if (ok) break;
    }
    return CCCS_OK;
}

struct BatchRef {
    int plan;   // index into the beam plans
    int first;  // first beamlet of the plan in this batch
    int n;
};

// Linear-in-depth bounds of a batch's footprint in the pillar frame, for the sample stage's cull.
// A sample is a footprint sample when its TERMA point lies in its beamlet's fluence cell, i.e.
// inside the pyramid from the source through the cell's corners at the isocentre plane
// (fluence_cell projects centrally from the source).  The pillar frame is a rotation of PVCS
// scaled by 1/rvs (isotropic spacing), so the pyramid stays a pyramid there, and its section at
// pillar depth y has the four corner rays' points as corners, whose x and z are linear in y.  The
// lowest corner x over the batch's beamlets is a minimum of linear functions (concave in y), the
// highest a maximum (convex), so the chords between the depths -0.5 and pd[1] - 0.5 bound them
// from below and above.  (A point on the beam's central axis has no defined cell and takes cell
// 0; the beamlet of cell 0 therefore also counts the axis line.)  out[0..3]: x lower and upper
// bounds of the local packed coordinate (pillar frame + 0.5) as c + s * py, py the packed depth
// coordinate, as {c_lo, s_lo, c_hi, s_hi}; out[4..7] the same for z; widened by `margin` samples.
void footprint_cull(const BatchDev& B, const BeamletDev* bl, v3 rvs, float margin, float out[8]) {
    const BeamDev& bm = B.beam;
    const double y0 = -0.5, y1 = (double)B.pd[1] - 0.5;
    double lo[2][2] = {{1e300, 1e300}, {1e300, 1e300}}, hi[2][2] = {{-1e300, -1e300}, {-1e300, -1e300}};  // [x|z][end]
    for (int q = 0; q < B.n; q++) {
        const BeamletDev& d = bl[q];
        // pillar frame of a PVCS point: fwd_m (p - pst) / rvs
        auto frame = [&](double px, double py, double pz, double f[3]) {
            const double r[3] = {px - d.pst.x, py - d.pst.y, pz - d.pst.z};
            for (int i = 0; i < 3; i++)
                f[i] = ((double)d.fwd_m.m[3 * i] * r[0] + (double)d.fwd_m.m[3 * i + 1] * r[1] +
                        (double)d.fwd_m.m[3 * i + 2] * r[2]) / (double)rvs.x;
        };
        double S[3];
        frame(bm.src.x, bm.src.y, bm.src.z, S);
        const int ix = d.fidx % bm.fmx, iz = d.fidx / bm.fmx;
        const double cx0 = (double)bm.src.x + (double)bm.iso_dist * bm.dir_n.x,
                     cy0 = (double)bm.src.y + (double)bm.iso_dist * bm.dir_n.y,
                     cz0 = (double)bm.src.z + (double)bm.iso_dist * bm.dir_n.z;
        const int ncorner = d.fidx == 0 ? 5 : 4;
        for (int k = 0; k < ncorner; k++) {
            v3 lat = mk3(0.f, 0.f, 0.f);
            if (k < 4) {
                const float ax = (float)(ix + (k & 1)) - 0.5f * (float)bm.fmx;
                const float az = (float)(iz + (k >> 1)) - 0.5f * (float)bm.fmy;
                lat = rot2(B.beam_inv, mk3(ax * bm.bsx, 0.f, az * bm.bsy));  // BEV -> PVCS
            }
            double C[3];
            frame(cx0 + lat.x, cy0 + lat.y, cz0 + lat.z, C);
            for (int e = 0; e < 2; e++) {
                const double y = e == 0 ? y0 : y1;
                const double u = (y - S[1]) / (C[1] - S[1]);
                const double x = S[0] + u * (C[0] - S[0]), z = S[2] + u * (C[2] - S[2]);
                lo[0][e] = std::min(lo[0][e], x); hi[0][e] = std::max(hi[0][e], x);
                lo[1][e] = std::min(lo[1][e], z); hi[1][e] = std::max(hi[1][e], z);
            }
        }
    }
    // local coordinate = frame + 0.5, and depth y = py - 0.5: bound(py) = v(y0) + 0.5 + slope * py
    for (int a = 0; a < 2; a++) {
        const double sl = (lo[a][1] - lo[a][0]) / (y1 - y0), sh = (hi[a][1] - hi[a][0]) / (y1 - y0);
        out[4 * a + 0] = (float)(lo[a][0] + 0.5 - margin);
        out[4 * a + 1] = (float)sl;
        out[4 * a + 2] = (float)(hi[a][0] + 0.5 + margin);
        out[4 * a + 3] = (float)sh;
    }
}

// Device tables and sizes of one job.
struct Job {
    std::vector<BatchDev> batches;
    std::vector<BeamletDev> beamlets;
    std::vector<BlConeDev> cones;          // [ray][batch]
    std::vector<uint64_t> ray_lines;       // lines per ray
    uint64_t max_lines = 0, max_rev = 0, pdose = 0, arena = 0, rows = 0;
    uint64_t pvox = 0;                     // pillar voxels of all beamlets (bounds the sparse output)
    uint32_t tiles = 0;
};

// Calc_bbox-relative voxel box of one pillar's unpacked region (rim excluded), widened by 2.
void pillar_box(const cccs_problem& p, const cccs_beam& b, const BeamletPlan& o, const int* pd, int* box0, int* boxn) {
    const v3 rvs = mk3(p.rev_latspacing, p.rev_longspacing, p.rev_latspacing);
    const Rot2 binv = rot2_beam_inv(b.azimuth, b.zenith, b.coll);
    const Rot1 r = beamlet_rot(o.theta, o.phi, true);
    const float lo = (float)(kWall + kBuffer) - 0.5f;
    double mn[3] = {1e30, 1e30, 1e30}, mx[3] = {-1e30, -1e30, -1e30};
    for (int c = 0; c < 8; c++) {
        v3 q = mk3((c & 1) ? (float)(pd[0] - kBuffer) - 0.5f : lo, (c & 2) ? (float)(pd[1] - kBuffer) - 0.5f : lo,
                   (c & 4) ? (float)(pd[2] - kBuffer) - 0.5f : lo);
        v3 g = rot2(binv, rot1(r, q * rvs)) + o.pst;
        for (int k = 0; k < 3; k++) {
            const double idx = ((double)comp(g, k) - p.start[k]) / p.voxel[k] - (double)p.bbox_start[k];
            mn[k] = std::min(mn[k], idx);
            mx[k] = std::max(mx[k], idx);
        }
    }
    for (int k = 0; k < 3; k++) {
        const long a = std::max(0L, (long)std::floor(mn[k]) - 2);
        const long e = std::min((long)p.bbox_size[k], (long)std::ceil(mx[k]) + 3);
        box0[k] = (int)a;
        boxn[k] = (int)std::max(0L, e - a);
    }
}

// A finished job's sparse columns in pinned host memory, and the event that marks them copied.
struct Delivery {
    bool live = false;
    hipEvent_t ev = nullptr;
    uint64_t* keys = nullptr;
    float* vals = nullptr;
    uint32_t* stat = nullptr;  // per column: kept, dropped, (two unused)
    uint32_t* rcnt = nullptr;  // per column and ROI: kept entries (ROI reduction)
    int nroi = 0;
    std::vector<std::pair<int, int>> cols;  // (request beam, fluence index) per column
};

int build_job(cccs_ctx* ctx, const std::vector<BeamPlan>& plans, const std::vector<BatchRef>& refs,
              const std::vector<const float*>& d_fmaps, Job& J) {
    const cccs_problem& p = ctx->prob;
    const v3 rvs = mk3(p.rev_latspacing, p.rev_longspacing, p.rev_latspacing);
    const int nb = (int)refs.size(), nrays = ctx->nrays;
    J = Job();
    J.batches.resize(nb);
    J.cones.resize((size_t)nrays * nb);
    J.ray_lines.assign(nrays, 0);
    std::vector<uint64_t> ray_rev(nrays, 0);
    uint32_t mr[3] = {p.max_rev_size[0], p.max_rev_size[1], p.max_rev_size[2]};
    if (mr[0] == 0) mr[0] = mr[1] = mr[2] = 500;
    for (int bi = 0; bi < nb; bi++) {
        const BatchRef& R = refs[bi];
        const BeamPlan& P = plans[R.plan];
        const cccs_beam& b = *P.b;
        BatchDev& B = J.batches[bi];
        B.beam = make_beam_dev(b);
        B.beam_inv = rot2_beam_inv(b.azimuth, b.zenith, b.coll);
        B.fmap = d_fmaps[R.plan];
        B.n = R.n;
        B.npx = (int)std::ceil(std::sqrt((float)R.n));
        B.npy = (int)std::ceil((float)R.n / (float)B.npx);
        for (int k = 0; k < 3; k++) B.pd[k] = P.pd[k];
        B.gd[0] = B.npx * P.pd[0] + kWall;
        B.gd[1] = P.pd[1] + kWall;
        B.gd[2] = B.npy * P.pd[2] + kWall;
        B.beamlet0 = (int)J.beamlets.size();
        B.pdose = J.pdose;
        J.pdose += (uint64_t)B.gd[0] * B.gd[1] * B.gd[2];
        B.tiles[0] = (B.gd[1] + 15) / 16;
        B.tiles[1] = (B.gd[2] + 3) / 4;
        B.tiles[2] = (B.gd[0] + kGatherX - 1) / kGatherX;
        B.tile0 = J.tiles;
        // footprint prefilter: the beamlet axis runs along the pillar's y axis through its x/z
        // centre; a sample farther from it than half the widest divergent beamlet (at the far end
        // of the pillar, with a 1.2x allowance for the tilt of the cell boundaries) plus one sample
        // cannot project into its beamlet's fluence cell
        B.fc[0] = 0.5f * (P.psize.x / rvs.x + (float)kWall);
        B.fc[1] = 0.5f * (P.psize.z / rvs.z + (float)kWall);
        B.fw = 0.5f * P.max_bs * 1.2f * (1.f + P.psize.y / P.min_exit) / std::min(rvs.x, rvs.z) + 1.f;
        if (std::fabs(rvs.x - rvs.y) > 1e-6f * rvs.x) B.fw = 1e30f;  // the bound assumes isotropic spacing
        const uint64_t nt = (uint64_t)J.tiles + (uint64_t)B.tiles[0] * B.tiles[1] * B.tiles[2];
        if (nt > 0xffffffffULL) return set_error(CCCS_EGEOM, "packed grids too large for one job");
        J.tiles = (uint32_t)nt;
        for (int q = 0; q < R.n; q++) {
            const BeamletPlan& o = P.bl[R.first + q];
            BeamletDev d;
            d.pst = o.pst;
            d.inv = beamlet_rot(o.theta, o.phi, true);
            d.fwd = beamlet_rot(o.theta, o.phi, false);
            d.fwd_m = to_f(dm_mul(dm_beamlet_fwd(o.theta, o.phi), dm_beam_fwd(b.azimuth, b.zenith, b.coll)));
            d.fidx = o.fidx;
            d.batch = bi;
            d.pix = q % B.npx;
            d.piy = q / B.npx;
            pillar_box(p, b, o, P.pd, d.box0, d.boxn);
            d.arena = J.arena;
            J.arena += (uint64_t)d.boxn[0] * d.boxn[1] * d.boxn[2];
            // row counts are 16-bit and row offsets 32-bit
            if (d.boxn[2] > 65535 || J.rows + (uint64_t)d.boxn[0] * d.boxn[1] > 0xffffffffULL)
                return set_error(CCCS_EGEOM, "column boxes too large for one job");
            d.row0 = (uint32_t)J.rows;
            J.rows += (uint64_t)d.boxn[0] * d.boxn[1];
            J.pvox += (uint64_t)P.pd[0] * P.pd[1] * P.pd[2];
            J.beamlets.push_back(d);
        }
        B.gw = B.fw;
        if (B.fw < 1e29f) {
            footprint_cull(B, J.beamlets.data() + B.beamlet0, rvs, 0.05f, B.cb);
            // the bounds' farthest reach from the axis (local frame: axis at fc + 0.5) at either end
            // of the grid's depth range; the bounds are linear, so the ends hold the maximum
            for (int e = 0; e < 2; e++) {
                const float py = e == 0 ? 0.f : (float)B.gd[1];
                for (int ax = 0; ax < 2; ax++) {
                    const float ctr = B.fc[ax] + 0.5f;
                    const float lo = B.cb[4 * ax] + B.cb[4 * ax + 1] * py, hi = B.cb[4 * ax + 2] + B.cb[4 * ax + 3] * py;
                    B.gw = e == 0 && ax == 0 ? std::max(ctr - lo, hi - ctr) : std::max(B.gw, std::max(ctr - lo, hi - ctr));
                }
            }
            B.gw = std::min(B.gw, B.fw);
        } else {  // anisotropic spacing: no bound
            const float nb_[8] = {-1e30f, 0.f, 1e30f, 0.f, -1e30f, 0.f, 1e30f, 0.f};
            std::copy(nb_, nb_ + 8, B.cb);
        }
        const v3 gsize = mk3((float)B.gd[0], (float)B.gd[1], (float)B.gd[2]) * rvs;
        for (int r = 0; r < nrays; r++) {
            const float th = ctx->theta[r] * 3.14159265358979323846f / 180.f;
            const float ph = ctx->phi[r] * 3.14159265358979323846f / 180.f;
            Rev rv;
            if (!find_rev(p, mr, gsize, th, ph, rv)) return set_error(CCCS_EGEOM, "REV grid exceeds max_rev_size");
            if (rv.size[0] > (uint32_t)kMaxLine) return set_error(CCCS_EGEOM, "REV line longer than 1024 samples");
            BlConeDev& c = J.cones[(size_t)r * nb + bi];
            c.kinv = rot2_kernel_inv(th, ph);
            c.kfwd = rot2_kernel_fwd(th, ph);
            c.kinv_m = to_f(dm_kernel_inv(th, ph));
            c.kfwd_m = to_f(dm_kernel_fwd(th, ph));
            c.minc = rv.minc;
            c.nx = (int)rv.size[0];
            c.ny = (int)rv.size[1];
            c.nz = (int)rv.size[2];
            c.batch = bi;
            c.line0 = J.ray_lines[r];
            J.ray_lines[r] += (uint64_t)c.ny * c.nz;
            ray_rev[r] += (uint64_t)c.nx * c.ny * c.nz;
        }
    }
    for (int r = 0; r < nrays; r++) {
        J.max_lines = std::max(J.max_lines, J.ray_lines[r]);
        J.max_rev = std::max(J.max_rev, ray_rev[r]);
    }
    if (J.max_lines > 0x7fffffffULL || J.max_rev > 0xffffffffULL)
        return set_error(CCCS_EGEOM, "REV grids of one cone exceed 2^32 samples in one job");
    return CCCS_OK;
}

BlArgs base_args(cccs_ctx* ctx) {
    const cccs_problem& p = ctx->prob;
    BlArgs a;
    std::memset(&a, 0, sizeof(a));
    a.dens = ctx->d_dens;
    a.tables = ctx->d_tables;
    a.ntheta = (int)p.ntheta;
    a.nradii = (int)p.nradii;
    a.nphi = (int)p.nphi;
    a.delr = p.rev_longspacing;
    a.rvs = mk3(p.rev_latspacing, p.rev_longspacing, p.rev_latspacing);
    a.rrv = mk3(1.f / a.rvs.x, 1.f / a.rvs.y, 1.f / a.rvs.z);
    a.vd = ctx->vol;
    a.spec = ctx->spec;
    a.bhc = p.beamhard_correct;
    return a;
}

inline unsigned grid_for(uint64_t n, unsigned block) { return (unsigned)std::min<uint64_t>((n + block - 1) / block, 1u << 20); }

// `while_busy` (may be empty) runs on the host once every launch of the job's first pass is
// enqueued, before the host waits for them: host work there overlaps the job on the device.
int run_job_cones(cccs_ctx* ctx, const Job& J, const std::function<int()>& while_busy) {
    const cccs_problem& p = ctx->prob;
    const int nb = (int)J.batches.size(), nrays = ctx->nrays;
    // device tables: batches | beamlets | cones
    const size_t sz_b = sizeof(BatchDev) * nb, sz_l = sizeof(BeamletDev) * J.beamlets.size(),
                 sz_c = sizeof(BlConeDev) * J.cones.size();
    const size_t o_l = (sz_b + 255) & ~size_t(255), o_c = (o_l + sz_l + 255) & ~size_t(255);
    const size_t total = o_c + sz_c;
    CCCS_TRY(ctx->bl_tab.ensure(total));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));  // the pinned staging buffer may still be in use
    CCCS_TRY(ctx->h_bl.ensure(total));
    uint8_t* h = ctx->h_bl.as<uint8_t>();
    std::memcpy(h, J.batches.data(), sz_b);
    std::memcpy(h + o_l, J.beamlets.data(), sz_l);
    std::memcpy(h + o_c, J.cones.data(), sz_c);
    uint8_t* d = ctx->bl_tab.as<uint8_t>();
    CCCS_HIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, ctx->stream));
    CCCS_TRY(ctx->bl_pdose.ensure(sizeof(float) * std::max<uint64_t>(1, J.pdose)));
    CCCS_TRY(ctx->lines.ensure(sizeof(LineMeta) * std::max<uint64_t>(1, J.max_lines)));
    CCCS_TRY(ctx->rev.ensure(sizeof(float) * std::max<uint64_t>(1, J.max_rev)));
    CCCS_TRY(ctx->bl_h.ensure(sizeof(float) * std::max<uint64_t>(1, J.max_rev)));
    CCCS_TRY(ctx->bl_qcnt.ensure(sizeof(uint32_t) * (size_t)std::max(1, nrays)));
    CCCS_TRY(ctx->h_small.ensure(sizeof(uint32_t) * kMaxRays + 64));
    // TERMA queue: footprint samples are a few percent of the stored samples; the counts are
    // checked after the pass and the pass reruns with larger queues if one overflowed
    uint64_t subcap = std::max<uint64_t>(uint64_t(1) << 14, J.max_rev / 8 / kSubQueues);
    CCCS_TRY(ctx->bl_qcnt.ensure(sizeof(uint32_t) * kSubQueues * kCntStride * (size_t)std::max(1, nrays)));
    CCCS_TRY(ctx->h_small.ensure(sizeof(uint32_t) * kSubQueues * kCntStride * kMaxRays));
    BlArgs a = base_args(ctx);
    a.batches = reinterpret_cast<const BatchDev*>(d);
    a.beamlets = reinterpret_cast<const BeamletDev*>(d + o_l);
    a.lines = ctx->lines.as<LineMeta>();
    a.rev = ctx->rev.as<float>();
    a.pdose = ctx->bl_pdose.as<float>();
    a.nbatch = nb;
    CCCS_TRY(ctx->bl_tiles.ensure(sizeof(uint32_t) * std::max<uint32_t>(1, J.tiles)));
    hipLaunchKernelGGL(bl_tile_batch_kernel, dim3(64, (unsigned)nb), dim3(256), 0, ctx->stream, a.batches,
                       ctx->bl_tiles.as<uint32_t>());
    CCCS_HIP(hipGetLastError());
    a.tile_batch = ctx->bl_tiles.as<uint32_t>();
    const int half = (int)p.ntheta / 2;
    // Several cone directions in flight (bl_streams = n): direction r runs on stream r % n with its
    // own line table, REV store, half steps, TERMA queue and scan partials, so consecutive
    // directions' sample, Siddon and walk overlap.  The gathers read-modify-write the shared packed
    // grids, so gather r waits for gather r-1 (an event).  A set is reused by direction r+n on the
    // same stream, after direction r's gather in stream order.
    const int nset = std::min(std::max(ctx->sched.bl_streams, 1), kRaySets);
    hipStream_t S[kRaySets];
    DevBuf *set_lines[kRaySets], *set_rev[kRaySets], *set_h[kRaySets], *set_q[kRaySets], *set_scan[kRaySets];
    S[0] = ctx->stream;
    set_lines[0] = &ctx->lines; set_rev[0] = &ctx->rev; set_h[0] = &ctx->bl_h; set_q[0] = &ctx->bl_queue;
    set_scan[0] = &ctx->scan;
    DevBuf* set_cone[kRaySets];
    set_cone[0] = &ctx->bl_lcone;
    for (int k = 1; k < kRaySets; k++) {
        S[k] = ctx->xstream[k - 1];
        set_lines[k] = &ctx->xlines[k - 1]; set_rev[k] = &ctx->xrev[k - 1]; set_h[k] = &ctx->xh[k - 1];
        set_q[k] = &ctx->xqueue[k - 1]; set_scan[k] = &ctx->xscan[k - 1]; set_cone[k] = &ctx->xlcone[k - 1];
    }
    for (int k = 0; k < nset; k++) CCCS_TRY(set_cone[k]->ensure(sizeof(uint32_t) * std::max<uint64_t>(1, J.max_lines)));
    for (int k = 1; k < nset; k++) {
        CCCS_TRY(set_lines[k]->ensure(sizeof(LineMeta) * std::max<uint64_t>(1, J.max_lines)));
        CCCS_TRY(set_rev[k]->ensure(sizeof(float) * std::max<uint64_t>(1, J.max_rev)));
        CCCS_TRY(set_h[k]->ensure(sizeof(float) * std::max<uint64_t>(1, J.max_rev)));
    }
    for (int attempt = 0; attempt < 3; attempt++) {
        subcap = std::min<uint64_t>(subcap, 0xffffffffULL / kSubQueues);
        for (int k = 0; k < nset; k++) CCCS_TRY(set_q[k]->ensure(sizeof(TermaReq) * subcap * kSubQueues));
        uint32_t* qcnt = ctx->bl_qcnt.as<uint32_t>();
        hipLaunchKernelGGL(bl_fill_kernel, dim3(grid_for(J.pdose, 256)), dim3(256), 0, ctx->stream, a.pdose, J.pdose, 0.f);
        hipLaunchKernelGGL(bl_fill_kernel, dim3(grid_for((uint64_t)kSubQueues * kCntStride * nrays, 256)), dim3(256), 0,
                           ctx->stream,
                           reinterpret_cast<float*>(qcnt), (uint64_t)kSubQueues * kCntStride * nrays, 0.f);
        CCCS_HIP(hipGetLastError());
        if (nset > 1) {  // the other streams start after the tables, fills and earlier work
            CCCS_HIP(hipEventRecord(ctx->ray_ev[kRaySets], S[0]));
            for (int k = 1; k < nset; k++) CCCS_HIP(hipStreamWaitEvent(S[k], ctx->ray_ev[kRaySets], 0));
        }
        int last = -1;  // set of the last gather enqueued
        int rc = CCCS_OK;
        for (int r = 0; r < nrays && rc == CCCS_OK; r++) {
            a.cones = reinterpret_cast<const BlConeDev*>(d + o_c) + (size_t)r * nb;
            a.nlines = J.ray_lines[r];
            a.rowf = r % half;
            a.rowb = (int)p.ntheta - 1 - (r % half);
            a.counts = nullptr;
            if (a.nlines == 0) continue;
            const int k = r % nset;
            // the helpers (line scan, profiling events) enqueue on ctx->stream and use ctx->scan
            ctx->stream = S[k];
            std::swap(ctx->scan, *set_scan[k]);
            a.lines = set_lines[k]->as<LineMeta>();
            a.line_cone = set_cone[k]->as<uint32_t>();
            a.rev = set_rev[k]->as<float>();
            float* H = set_h[k]->as<float>();
            TermaReq* q = set_q[k]->as<TermaReq>();
            do {
                hipLaunchKernelGGL(bl_interval_kernel, dim3((unsigned)((a.nlines + 255) / 256)), dim3(256), 0,
                                   ctx->stream, a);
                if (hipGetLastError() != hipSuccess) { rc = set_error(CCCS_EDEVICE, "interval kernel"); break; }
                if ((rc = launch_line_scan(ctx, a.lines, a.nlines)) != CCCS_OK) break;
                uint32_t maxn = 1;
                for (int bi = 0; bi < nb; bi++) maxn = std::max(maxn, (uint32_t)J.cones[(size_t)r * nb + bi].nx);
                const unsigned lw = (unsigned)std::min<uint64_t>(a.nlines, (uint64_t)ctx->sched.bl_sample_waves);
                const unsigned ww = (unsigned)std::min<uint64_t>(a.nlines, (uint64_t)ctx->sched.bl_walk_waves);
                a.counts = prof_counter(ctx, lw);
                int tok = prof_begin(ctx, K_BL_SAMPLE);
                hipLaunchKernelGGL(bl_sample_kernel, dim3(lw), dim3(64), 0, ctx->stream, a, H, q, (uint32_t)subcap,
                                   qcnt + (size_t)kSubQueues * kCntStride * r);
                prof_end(ctx, tok, 0.0, 0.0, a.nlines);
                if (hipGetLastError() != hipSuccess) { rc = set_error(CCCS_EDEVICE, "sample kernel"); break; }
                {
                    const unsigned sb = kSubQueues * 64;
                    uint32_t* cnt = prof_counter(ctx, sb);
                    int ts = prof_begin(ctx, K_BL_SIDDON);
                    hipLaunchKernelGGL(bl_siddon_kernel, dim3(sb), dim3(256), 0, ctx->stream, a, q, (uint32_t)subcap,
                                       (const uint32_t*)(qcnt + (size_t)kSubQueues * kCntStride * r), cnt);
                    prof_end(ctx, ts, 0.0, 0.0, 0);
                    if (hipGetLastError() != hipSuccess) { rc = set_error(CCCS_EDEVICE, "siddon kernel"); break; }
                }
                // length classes (bl_walk_split, twice that, ...; the last ends at the longest
                // line): each launch sizes its LDS for its own longest line, so the short lines,
                // the bulk after trimming, run with more waves per CU
                const uint32_t split = (uint32_t)ctx->sched.bl_walk_split;
                uint32_t bounds[12];
                int nclass = 0;
                for (uint32_t hi = split; split > 0 && hi < maxn && nclass < 11; hi *= 2) bounds[nclass++] = hi;
                bounds[nclass++] = maxn;
                const size_t per = sizeof(double) + sizeof(float2) + 3 * sizeof(short);
                uint32_t* wc = prof_counter(ctx, (size_t)nclass * ww);
                int tw = prof_begin(ctx, K_BL_WALK);
                for (int ci = 0; ci < nclass; ci++) {
                    a.counts = wc ? wc + (size_t)ci * ww : nullptr;
                    hipLaunchKernelGGL(bl_walk_kernel, dim3(ww), dim3(64), per * bounds[ci], ctx->stream, a,
                                       (const float*)H, (int)bounds[ci], ci == 0 ? 1 : (int)bounds[ci - 1] + 1);
                }
                prof_end(ctx, tw, 0.0, 0.0, a.nlines);
                if (hipGetLastError() != hipSuccess) { rc = set_error(CCCS_EDEVICE, "walk kernel"); break; }
                a.counts = nullptr;
                if (nset > 1 && last >= 0 && last != k &&
                    hipStreamWaitEvent(ctx->stream, ctx->ray_ev[last], 0) != hipSuccess) {
                    rc = set_error(CCCS_EDEVICE, "stream wait");
                    break;
                }
                int tk = prof_begin(ctx, K_BL_GATHER);
                hipLaunchKernelGGL(bl_gather_kernel, dim3(J.tiles), dim3(64), 0, ctx->stream, a, J.tiles);
                prof_end(ctx, tk, 8.0 * (double)J.pdose, 0.0, J.pdose);
                if (hipGetLastError() != hipSuccess) { rc = set_error(CCCS_EDEVICE, "gather kernel"); break; }
                if (nset > 1 && hipEventRecord(ctx->ray_ev[k], ctx->stream) != hipSuccess) {
                    rc = set_error(CCCS_EDEVICE, "event");
                    break;
                }
                last = k;
            } while (false);
            std::swap(ctx->scan, *set_scan[k]);
            ctx->stream = S[0];
        }
        // join: the first stream continues after every direction's work (the gathers are chained,
        // and no stream has work after its last gather)
        if (nset > 1 && last > 0) CCCS_HIP(hipStreamWaitEvent(S[0], ctx->ray_ev[last], 0));
        if (rc != CCCS_OK) {
            for (int k = 0; k < nset; k++) (void)hipStreamSynchronize(S[k]);
            return rc;
        }
        if (attempt == 0 && while_busy) {
            const int hr = while_busy();
            if (hr != CCCS_OK) {
                for (int k = 0; k < nset; k++) (void)hipStreamSynchronize(S[k]);
                return hr;
            }
        }
        uint32_t* hc = ctx->h_small.as<uint32_t>();
        CCCS_HIP(hipMemcpyAsync(hc, qcnt, sizeof(uint32_t) * kSubQueues * kCntStride * nrays, hipMemcpyDeviceToHost,
                                ctx->stream));
        CCCS_HIP(hipStreamSynchronize(ctx->stream));
        uint64_t need = 0;
        for (int r = 0; r < kSubQueues * nrays; r++) need = std::max<uint64_t>(need, hc[(size_t)r * kCntStride]);
        if (need <= subcap) return CCCS_OK;
        subcap = need + need / 4;
    }
    return set_error(CCCS_ENOMEM, "TERMA queue overflow");
}

int run_job_unpack(cccs_ctx* ctx, const Job& J) {
    const cccs_problem& p = ctx->prob;
    CCCS_TRY(ctx->bl_arena.ensure(sizeof(float) * std::max<uint64_t>(1, J.arena)));
    const size_t sz_b = sizeof(BatchDev) * J.batches.size();
    const size_t o_l = (sz_b + 255) & ~size_t(255);
    UnpackArgs u;
    u.batches = ctx->bl_tab.as<BatchDev>();
    u.beamlets = reinterpret_cast<const BeamletDev*>(ctx->bl_tab.as<uint8_t>() + o_l);
    u.pdose = ctx->bl_pdose.as<float>();
    u.arena = ctx->bl_arena.as<float>();
    u.nbeamlets = (int)J.beamlets.size();
    u.narena = J.arena;
    u.start = ld3(p.start);
    u.voxel = ld3(p.voxel);
    u.rvs = mk3(p.rev_latspacing, p.rev_longspacing, p.rev_latspacing);
    for (int k = 0; k < 3; k++) u.bbs[k] = (int)p.bbox_start[k];
    if (J.arena == 0) return CCCS_OK;
    const uint64_t blocks = (J.arena + 255) / 256;
    if (blocks > 0x7fffffffULL) return set_error(CCCS_EGEOM, "column arena too large for one job");
    int tok = prof_begin(ctx, K_BL_UNPACK);
    hipLaunchKernelGGL(bl_unpack_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, u);
    prof_end(ctx, tok, 8.0 * (double)J.arena, 0.0, J.arena);
    CCCS_HIP(hipGetLastError());
    return CCCS_OK;
}

// Keeps v > t * max(column) in the reference's emission order (_sparsify, sparse_io.cpp:22-58) for a
// column stored i outer, k inner over the whole calc_bbox.
void sparsify_host(const float* col, const cccs_problem& p, float thresh, std::vector<uint64_t>& keys,
                   std::vector<float>& vals, uint64_t& dropped) {
    const uint64_t bx = p.bbox_size[0], by = p.bbox_size[1], bz = p.bbox_size[2], n = bx * by * bz;
    float mx = 0.f;
    for (uint64_t e = 0; e < n; e++) mx = std::max(mx, col[e]);
    const float cut = thresh * mx;
    keys.clear();
    vals.clear();
    dropped = 0;
    uint64_t e = 0;
    for (uint64_t i = 0; i < bx; i++)
        for (uint64_t j = 0; j < by; j++)
            for (uint64_t k = 0; k < bz; k++, e++) {
                const float v = col[e];
                if (v <= cut) {
                    if (v > 0.f) dropped++;
                    continue;
                }
                keys.push_back(((k + p.bbox_start[2]) * p.size[1] + (j + p.bbox_start[1])) * p.size[0] +
                               (i + p.bbox_start[0]));
                vals.push_back(v);
            }
}

// The plain path for one batch (see plain_ray_kernel): per cone raytrace / convolve / gather
// launches, then per beamlet a full calc_bbox unpack copied to the host and sparsified there by
// worker threads while the device unpacks the next group.
int run_plain_batch(cccs_ctx* ctx, const std::vector<BeamPlan>& plans, const BatchRef& R,
                    const std::vector<const float*>& d_fmaps, float thresh, cccs_sparse_sink sink, void* user) {
    const cccs_problem& p = ctx->prob;
    Job J;
    CCCS_TRY(build_job(ctx, plans, std::vector<BatchRef>{R}, d_fmaps, J));
    const uint64_t nbb = (uint64_t)p.bbox_size[0] * p.bbox_size[1] * p.bbox_size[2];
    for (BeamletDev& d : J.beamlets) {
        for (int k = 0; k < 3; k++) { d.box0[k] = 0; d.boxn[k] = (int)p.bbox_size[k]; }
        d.arena = 0;
    }
    J.arena = nbb;
    const int nrays = ctx->nrays;
    const size_t sz_b = sizeof(BatchDev), sz_l = sizeof(BeamletDev) * J.beamlets.size(),
                 sz_c = sizeof(BlConeDev) * J.cones.size();
    const size_t o_l = (sz_b + 255) & ~size_t(255), o_c = (o_l + sz_l + 255) & ~size_t(255);
    CCCS_TRY(ctx->bl_tab.ensure(o_c + sz_c));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_TRY(ctx->h_bl.ensure(o_c + sz_c));
    uint8_t* h = ctx->h_bl.as<uint8_t>();
    std::memcpy(h, J.batches.data(), sz_b);
    std::memcpy(h + o_l, J.beamlets.data(), sz_l);
    std::memcpy(h + o_c, J.cones.data(), sz_c);
    uint8_t* d = ctx->bl_tab.as<uint8_t>();
    CCCS_HIP(hipMemcpyAsync(d, h, o_c + sz_c, hipMemcpyHostToDevice, ctx->stream));
    uint64_t max_rev = 0, max_lines = 0;
    for (const BlConeDev& c : J.cones) {
        max_rev = std::max<uint64_t>(max_rev, (uint64_t)c.nx * c.ny * c.nz);
        max_lines = std::max<uint64_t>(max_lines, (uint64_t)c.ny * c.nz);
    }
    if (max_rev > 0xffffffffULL) return set_error(CCCS_EGEOM, "REV grid exceeds 2^32 samples");
    CCCS_TRY(ctx->bl_pdose.ensure(sizeof(float) * std::max<uint64_t>(1, J.pdose)));
    CCCS_TRY(ctx->lines.ensure(sizeof(LineMeta) * std::max<uint64_t>(1, max_lines)));
    CCCS_TRY(ctx->rev.ensure(sizeof(float) * std::max<uint64_t>(1, max_rev)));
    CCCS_TRY(ctx->bl_h.ensure(sizeof(float) * 2 * std::max<uint64_t>(1, max_rev)));
    BlArgs a = base_args(ctx);
    a.batches = reinterpret_cast<const BatchDev*>(d);
    a.beamlets = reinterpret_cast<const BeamletDev*>(d + o_l);
    a.lines = ctx->lines.as<LineMeta>();
    a.rev = ctx->rev.as<float>();
    a.pdose = ctx->bl_pdose.as<float>();
    a.nbatch = 1;
    float* dd = ctx->bl_h.as<float>();
    float* dt = dd + max_rev;
    hipLaunchKernelGGL(bl_fill_kernel, dim3(grid_for(J.pdose, 256)), dim3(256), 0, ctx->stream, a.pdose, J.pdose, 0.f);
    const int half = (int)p.ntheta / 2;
    for (int r = 0; r < nrays; r++) {
        const BlConeDev& c = J.cones[r];
        a.cones = reinterpret_cast<const BlConeDev*>(d + o_c) + r;
        a.nlines = (uint64_t)c.ny * c.nz;
        a.rowf = r % half;
        a.rowb = (int)p.ntheta - 1 - (r % half);
        const uint64_t nrev = (uint64_t)c.nx * c.ny * c.nz;
        if (nrev == 0) continue;
        int t0 = prof_begin(ctx, K_PL_RAY);
        hipLaunchKernelGGL(plain_ray_kernel, dim3((unsigned)((nrev + 255) / 256)), dim3(256), 0, ctx->stream, a, dd, dt);
        prof_end(ctx, t0, 8.0 * (double)nrev, 0.0, nrev);
        int t1 = prof_begin(ctx, K_PL_CONV);
        hipLaunchKernelGGL(plain_conv_kernel, dim3((unsigned)a.nlines), dim3(256), sizeof(float) * 2 * c.nx, ctx->stream,
                           a, (const float*)dd, (const float*)dt, a.rev);
        prof_end(ctx, t1, 12.0 * (double)nrev, 0.0, nrev);
        hipLaunchKernelGGL(plain_meta_kernel, dim3((unsigned)((a.nlines + 255) / 256)), dim3(256), 0, ctx->stream, a.lines,
                           a.nlines, c.nx);
        int t2 = prof_begin(ctx, K_BL_GATHER);
        hipLaunchKernelGGL(bl_gather_kernel, dim3(J.tiles), dim3(64), 0, ctx->stream, a, J.tiles);
        prof_end(ctx, t2, 8.0 * (double)J.pdose, 0.0, J.pdose);
        CCCS_HIP(hipGetLastError());
    }
    // per beamlet: dense unpack -> host -> sparsify (groups of G, two pinned slot sets)
    const int G = 8, nbl = (int)J.beamlets.size();
    CCCS_TRY(ctx->bl_arena.ensure(sizeof(float) * nbb * G));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_TRY(ctx->h_bl_out.ensure(sizeof(float) * nbb * 2 * G));
    float* pin = ctx->h_bl_out.as<float>();
    std::vector<std::vector<uint64_t>> keys(2 * G);
    std::vector<std::vector<float>> vals(2 * G);
    std::vector<uint64_t> drop(2 * G);
    std::vector<std::thread> workers;
    struct Joiner {  // no early return leaves a worker running
        std::vector<std::thread>& w;
        ~Joiner() {
            for (std::thread& t : w)
                if (t.joinable()) t.join();
        }
    } joiner{workers};
    int pending0 = -1;  // first beamlet of the group being sparsified
    auto emit = [&](int g0) -> int {
        for (std::thread& t : workers) t.join();
        workers.clear();
        const int set = (g0 / G) % 2;
        for (int q = g0; q < std::min(nbl, g0 + G); q++) {
            const int sl = set * G + (q - g0);
            cccs_sparse_column col{};
            col.beam_index = plans[R.plan].beam;
            col.beamlet_index = J.beamlets[q].fidx;
            col.nnz = keys[sl].size();
            col.keys = keys[sl].data();
            col.values = vals[sl].data();
            col.perc_nonzero = (float)(100.0 * (double)keys[sl].size() / (double)nbb);
            col.perc_dropped = (float)(100.0 * (double)drop[sl] / (double)nbb);
            if (sink(user, &col) != 0) return set_error(CCCS_EIO, "sparse sink aborted the call");
        }
        return CCCS_OK;
    };
    UnpackArgs u;
    u.batches = reinterpret_cast<const BatchDev*>(d);
    u.pdose = a.pdose;
    u.nbeamlets = 1;
    u.narena = nbb;
    u.start = ld3(p.start);
    u.voxel = ld3(p.voxel);
    u.rvs = a.rvs;
    for (int k = 0; k < 3; k++) u.bbs[k] = (int)p.bbox_start[k];
    for (int g0 = 0; g0 < nbl; g0 += G) {
        const int set = (g0 / G) % 2;
        for (int q = g0; q < std::min(nbl, g0 + G); q++) {
            u.beamlets = reinterpret_cast<const BeamletDev*>(d + o_l) + q;
            u.arena = ctx->bl_arena.as<float>() + nbb * (size_t)(q - g0);
            int tk = prof_begin(ctx, K_BL_UNPACK);
            hipLaunchKernelGGL(bl_unpack_kernel, dim3((unsigned)((nbb + 255) / 256)), dim3(256), 0, ctx->stream, u);
            prof_end(ctx, tk, 8.0 * (double)nbb, 0.0, nbb);
            CCCS_HIP(hipMemcpyAsync(pin + nbb * (size_t)(set * G + (q - g0)), u.arena, sizeof(float) * nbb,
                                    hipMemcpyDeviceToHost, ctx->stream));
        }
        CCCS_HIP(hipStreamSynchronize(ctx->stream));
        if (pending0 >= 0) CCCS_TRY(emit(pending0));
        for (int q = g0; q < std::min(nbl, g0 + G); q++) {
            const int sl = set * G + (q - g0);
            const float* col = pin + nbb * (size_t)sl;
            workers.emplace_back([&, col, sl] { sparsify_host(col, p, thresh, keys[sl], vals[sl], drop[sl]); });
        }
        pending0 = g0;
    }
    if (pending0 >= 0) CCCS_TRY(emit(pending0));
    return CCCS_OK;
}

// One batch's share of a job's device footprint: REV samples and lines per cone direction, and
// the packed grid plus the sparse-output bound (12 B per pillar voxel) in bytes.
int batch_footprint(const cccs_ctx* ctx, const std::vector<BeamPlan>& plans, const BatchRef& R,
                    std::vector<uint64_t>& rev, std::vector<uint64_t>& lines, uint64_t& fixed) {
    const cccs_problem& p = ctx->prob;
    const BeamPlan& P = plans[R.plan];
    const v3 rvs = mk3(p.rev_latspacing, p.rev_longspacing, p.rev_latspacing);
    uint32_t mr[3] = {p.max_rev_size[0], p.max_rev_size[1], p.max_rev_size[2]};
    if (mr[0] == 0) mr[0] = mr[1] = mr[2] = 500;
    const int npx = (int)std::ceil(std::sqrt((float)R.n));
    const int npy = (int)std::ceil((float)R.n / (float)npx);
    const uint64_t gd0 = (uint64_t)npx * P.pd[0] + kWall, gd1 = (uint64_t)P.pd[1] + kWall,
                   gd2 = (uint64_t)npy * P.pd[2] + kWall;
    const v3 gsize = mk3((float)gd0, (float)gd1, (float)gd2) * rvs;
    rev.assign(ctx->nrays, 0);
    lines.assign(ctx->nrays, 0);
    for (int r = 0; r < ctx->nrays; r++) {
        const float th = ctx->theta[r] * 3.14159265358979323846f / 180.f;
        const float ph = ctx->phi[r] * 3.14159265358979323846f / 180.f;
        Rev rv;
        if (!find_rev(p, mr, gsize, th, ph, rv)) return set_error(CCCS_EGEOM, "REV grid exceeds max_rev_size");
        rev[r] = (uint64_t)rv.size[0] * rv.size[1] * rv.size[2];
        lines[r] = (uint64_t)rv.size[1] * rv.size[2];
    }
    fixed = sizeof(float) * gd0 * gd1 * gd2 + 12ull * (uint64_t)R.n * P.pd[0] * P.pd[1] * P.pd[2];
    return CCCS_OK;
}


}  // namespace

}  // namespace cccs

using namespace cccs;

// =================================================================================================
// C-ABI

namespace {

// Uploads the fluence maps of the plans (one device block, pinned staging).
int upload_fmaps(cccs_ctx* ctx, const std::vector<BeamPlan>& plans, const float* const* fluence,
                 std::vector<const float*>& d_fmaps) {
    size_t total = 0;
    std::vector<size_t> off(plans.size());
    for (size_t i = 0; i < plans.size(); i++) {
        off[i] = total;
        total += (size_t)plans[i].b->fmap_size[0] * plans[i].b->fmap_size[1];
    }
    CCCS_TRY(ctx->bl_fmap.ensure(sizeof(float) * std::max<size_t>(1, total)));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_TRY(ctx->h_bl_out.ensure(sizeof(float) * std::max<size_t>(1, total)));
    float* h = ctx->h_bl_out.as<float>();
    for (size_t i = 0; i < plans.size(); i++)
        std::memcpy(h + off[i], fluence[plans[i].beam],
                    sizeof(float) * plans[i].b->fmap_size[0] * plans[i].b->fmap_size[1]);
    CCCS_HIP(hipMemcpyAsync(ctx->bl_fmap.p, h, sizeof(float) * total, hipMemcpyHostToDevice, ctx->stream));
    d_fmaps.resize(plans.size());
    for (size_t i = 0; i < plans.size(); i++) d_fmaps[i] = ctx->bl_fmap.as<float>() + off[i];
    return CCCS_OK;
}

// Runs the cone passes and the unpack of one job (the arena then holds every beamlet's box).
int run_job(cccs_ctx* ctx, const std::vector<BeamPlan>& plans, const std::vector<BatchRef>& refs,
            const std::vector<const float*>& d_fmaps, Job& J) {
    CCCS_TRY(build_job(ctx, plans, refs, d_fmaps, J));
    CCCS_TRY(run_job_cones(ctx, J, std::function<int()>()));
    return run_job_unpack(ctx, J);
}

std::vector<BatchRef> all_batches(const std::vector<BeamPlan>& plans) {
    std::vector<BatchRef> refs;
    for (size_t i = 0; i < plans.size(); i++) {
        int first = 0;
        for (int n : plans[i].bsizes) {
            refs.push_back(BatchRef{(int)i, first, n});
            first += n;
        }
    }
    return refs;
}

}  // namespace

namespace {

// ---- ROI reduction (_reduce_by_roi, DoseCalcIO/sparse_io.cpp:62-158) ------------------------------
constexpr int kMaxRoi = 64;

struct RoiDev {
    int c0[3], cs[3];    // crop box start and size on the CT grid
    uint64_t rank_off;   // first rank entry (crop box, x fastest; -1 outside the mask)
    uint64_t mark;       // voxels of all earlier ROIs (first key of this block)
};

struct RoiArgs {
    const RoiDev* rois;
    const int32_t* rank;
    int nroi;
    int nx, ny;                  // CT grid
    const uint64_t* keys;        // the job's kept entries, column after column
    const float* vals;
    const uint64_t* off;         // first entry of each column (nbl + 1)
    uint32_t* cnt;               // [column][roi] reduced counts
    const uint64_t* roff;        // [column][roi] first reduced entry (exclusive scan of cnt)
    uint64_t* rkeys;
    float* rvals;
};

// ROI rank of CT voxel `key` in ROI r (or -1).
__device__ __forceinline__ int32_t roi_rank(const RoiArgs& a, const RoiDev& R, int x, int y, int z) {
    const int lx = x - R.c0[0], ly = y - R.c0[1], lz = z - R.c0[2];
    if (lx < 0 || ly < 0 || lz < 0 || lx >= R.cs[0] || ly >= R.cs[1] || lz >= R.cs[2]) return -1;
    return a.rank[R.rank_off + (uint64_t)lx + (uint64_t)R.cs[0] * ((uint64_t)ly + (uint64_t)R.cs[1] * lz)];
}

// Pass 1: per column and ROI, the number of kept entries inside the ROI (one LDS add per wave,
// chunk and ROI).
__global__ void __launch_bounds__(256) roi_count_kernel(RoiArgs a) {
    __shared__ uint32_t s_c[kMaxRoi];
    const int q = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    for (int r = tid; r < a.nroi; r += 256) s_c[r] = 0u;
    __syncthreads();
    const uint64_t e0 = a.off[q], e1 = a.off[q + 1];
    const uint64_t plane = (uint64_t)a.nx * (uint64_t)a.ny;
    for (uint64_t base = e0; base < e1; base += 256) {
        const uint64_t e = base + tid;
        int x = -1, y = -1, z = -1;
        if (e < e1) {
            const uint64_t key = a.keys[e];
            z = (int)(key / plane);
            const int rem = (int)(key - (uint64_t)z * plane);
            y = rem / a.nx;
            x = rem - y * a.nx;
        }
        for (int r = 0; r < a.nroi; r++) {
            const uint64_t m = __ballot(e < e1 && roi_rank(a, a.rois[r], x, y, z) >= 0);
            if (lane == 0 && m) atomicAdd(&s_c[r], (uint32_t)__popcll(m));
        }
    }
    __syncthreads();
    for (int r = tid; r < a.nroi; r += 256) a.cnt[(size_t)q * a.nroi + r] = s_c[r];
}

// Pass 2: ROI by ROI, the column's entries inside the ROI in the column's order (an
// order-preserving compaction per 256-entry chunk), keyed mark + rank.
__global__ void __launch_bounds__(256) roi_emit_kernel(RoiArgs a) {
    __shared__ uint32_t s_w[4];
    const int q = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint64_t e0 = a.off[q], e1 = a.off[q + 1];
    const uint64_t plane = (uint64_t)a.nx * (uint64_t)a.ny;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int r = 0; r < a.nroi; r++) {
        const RoiDev R = a.rois[r];
        uint64_t pos = a.roff[(size_t)q * a.nroi + r];
        for (uint64_t base = e0; base < e1; base += 256) {
            const uint64_t e = base + tid;
            int32_t rk = -1;
            if (e < e1) {
                const uint64_t key = a.keys[e];
                const int z = (int)(key / plane), rem = (int)(key - (uint64_t)z * plane);
                const int y = rem / a.nx, x = rem - y * a.nx;
                rk = roi_rank(a, R, x, y, z);
            }
            const uint64_t m = __ballot(rk >= 0);
            if (lane == 0) s_w[wid] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t before = 0, total = 0;
            for (int w = 0; w < 4; w++) {
                before += w < wid ? s_w[w] : 0u;
                total += s_w[w];
            }
            if (rk >= 0) {
                const uint64_t o = pos + before + (uint32_t)__popcll(m & lt);
                a.rkeys[o] = R.mark + (uint64_t)rk;
                a.rvals[o] = a.vals[e];
            }
            pos += total;
            __syncthreads();
        }
    }
}

// Exclusive scan of n uint32 counts into n+1 uint64 offsets (one workgroup, serial chunks).
__global__ void __launch_bounds__(1024) roi_scan_kernel(const uint32_t* __restrict__ cnt, size_t n,
                                                        uint64_t* __restrict__ out) {
    __shared__ uint64_t s_p[1024];
    const int tid = threadIdx.x;
    const size_t per = (n + 1023) / 1024;
    const size_t c0 = std::min(n, (size_t)tid * per), c1 = std::min(n, c0 + per);
    uint64_t loc = 0;
    for (size_t i = c0; i < c1; i++) loc += cnt[i];
    s_p[tid] = loc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint64_t v = tid >= o ? s_p[tid - o] : 0;
        __syncthreads();
        s_p[tid] += v;
        __syncthreads();
    }
    uint64_t run = s_p[tid] - loc;
    for (size_t i = c0; i < c1; i++) {
        out[i] = run;
        run += cnt[i];
    }
    if (tid == 1023) out[n] = s_p[1023];
}

}  // namespace

extern "C" int cccs_set_roi_masks(cccs_ctx* ctx, int nroi, const uint8_t* const* masks, const uint32_t* crop_start,
                                  const uint32_t* crop_size) {
    clear_error();
    if (!ctx || nroi < 0 || (nroi > 0 && (!masks || !crop_start || !crop_size))) return set_error(CCCS_EINVAL, "null argument");
    if (nroi > kMaxRoi) return set_error(CCCS_EINVAL, "at most " + std::to_string(kMaxRoi) + " ROIs");
    CCCS_HIP(hipSetDevice(ctx->device));
    ctx->nroi = 0;
    if (nroi == 0) return CCCS_OK;
    const cccs_problem& p = ctx->prob;
    std::vector<RoiDev> tab((size_t)nroi);
    std::vector<int32_t> rank;
    uint64_t mark = 0;
    for (int r = 0; r < nroi; r++) {
        RoiDev& R = tab[(size_t)r];
        for (int k = 0; k < 3; k++) {
            R.c0[k] = (int)crop_start[3 * r + k];
            R.cs[k] = (int)crop_size[3 * r + k];
            if (crop_start[3 * r + k] + (uint64_t)crop_size[3 * r + k] > p.size[k])
                return set_error(CCCS_EINVAL, "ROI crop box outside the CT grid");  // sparse_io.cpp:111-113
        }
        if (!masks[r]) return set_error(CCCS_EINVAL, "null ROI mask");
        R.rank_off = rank.size();
        R.mark = mark;
        const size_t n = (size_t)R.cs[0] * R.cs[1] * R.cs[2];
        rank.resize(rank.size() + n, -1);
        int32_t* rk = rank.data() + R.rank_off;
        int32_t elem = 0;
        // the reference's iteration: ii (x) outer, jj (y), kk (z) inner over the crop box
        for (int ii = 0; ii < R.cs[0]; ii++)
            for (int jj = 0; jj < R.cs[1]; jj++)
                for (int kk = 0; kk < R.cs[2]; kk++) {
                    const size_t ck = ((size_t)kk * R.cs[1] + jj) * R.cs[0] + ii;
                    if (masks[r][ck]) rk[ck] = elem++;
                }
        mark += (uint64_t)elem;
    }
    CCCS_TRY(ctx->roi_tab.ensure(sizeof(RoiDev) * tab.size()));
    CCCS_TRY(ctx->roi_rank.ensure(sizeof(int32_t) * std::max<size_t>(1, rank.size())));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_HIP(hipMemcpy(ctx->roi_tab.p, tab.data(), sizeof(RoiDev) * tab.size(), hipMemcpyHostToDevice));
    if (!rank.empty())
        CCCS_HIP(hipMemcpy(ctx->roi_rank.p, rank.data(), sizeof(int32_t) * rank.size(), hipMemcpyHostToDevice));
    ctx->nroi = nroi;
    return CCCS_OK;
}

extern "C" int cccs_beamlet_footprint(const cccs_ctx* ctx, const cccs_beam* beam, const float* fluence, uint64_t* bytes) {
    clear_error();
    if (!ctx || !beam || !fluence || !bytes) return set_error(CCCS_EINVAL, "null argument");
    *bytes = 0;
    CCCS_TRY(check_beam(ctx, *beam));
    std::vector<BeamPlan> plans(1);
    CCCS_TRY(plan_beam(ctx, *beam, fluence, 0, plans[0]));
    if (plans[0].bl.empty()) return CCCS_OK;
    const std::vector<BatchRef> refs = all_batches(plans);
    std::vector<uint64_t> rev(ctx->nrays, 0), lines(ctx->nrays, 0), r1, l1;
    uint64_t fixed = 0, f1 = 0;
    for (const BatchRef& R : refs) {
        CCCS_TRY(batch_footprint(ctx, plans, R, r1, l1, f1));
        for (int r = 0; r < ctx->nrays; r++) { rev[r] += r1[r]; lines[r] += l1[r]; }
        fixed += f1;
    }
    uint64_t mr = 0, ml = 0;
    for (int r = 0; r < ctx->nrays; r++) { mr = std::max(mr, rev[r]); ml = std::max(ml, lines[r]); }
    *bytes = fixed + sizeof(float) * 2 * mr + sizeof(LineMeta) * ml;
    return CCCS_OK;
}

extern "C" int cccs_sink_count(void* user, const cccs_sparse_column* column) {
    if (!user || !column) return 1;
    cccs_sink_totals* t = static_cast<cccs_sink_totals*>(user);
    t->columns++;
    t->nnz += column->nnz;
    double s = 0.0;
    for (uint64_t i = 0; i < column->nnz; i++) s += column->values[i];
    t->value_sum += s;
    return 0;
}

extern "C" int cccs_beamlet_dense(cccs_ctx* ctx, const cccs_beam* beam, const float* fluence, int mode, float* out,
                                  int max_beamlets, int* nactive) {
    clear_error();
    if (!ctx || !beam || !fluence || !out) return set_error(CCCS_EINVAL, "null argument");
    if (mode != CCCS_BEAMLET_FAITHFUL) return set_error(CCCS_EINVAL, "unknown beamlet mode");
    CCCS_HIP(hipSetDevice(ctx->device));
    CCCS_TRY(check_beam(ctx, *beam));
    std::vector<BeamPlan> plans(1);
    CCCS_TRY(plan_beam(ctx, *beam, fluence, 0, plans[0]));
    const int n = (int)plans[0].bl.size();
    if (nactive) *nactive = n;
    if (n == 0) return CCCS_OK;
    if (n > max_beamlets) return set_error(CCCS_EINVAL, "more active beamlets than max_beamlets");
    const cccs_problem& p = ctx->prob;
    const size_t nbb = (size_t)p.bbox_size[0] * p.bbox_size[1] * p.bbox_size[2];
    std::memset(out, 0, sizeof(float) * nbb * n);
    std::vector<const float*> d_fmaps;
    const float* fl[1] = {fluence};
    CCCS_TRY(upload_fmaps(ctx, plans, fl, d_fmaps));
    // one job per batch, as the reference processes its batches
    int done = 0;
    for (const BatchRef& R : all_batches(plans)) {
        Job J;
        CCCS_TRY(run_job(ctx, plans, std::vector<BatchRef>{R}, d_fmaps, J));
        std::vector<float> arena(J.arena);
        CCCS_HIP(hipStreamSynchronize(ctx->stream));
        if (J.arena) CCCS_HIP(hipMemcpy(arena.data(), ctx->bl_arena.p, sizeof(float) * J.arena, hipMemcpyDeviceToHost));
        for (int q = 0; q < R.n; q++) {
            const BeamletDev& d = J.beamlets[q];
            float* col = out + (size_t)(done + q) * nbb;
            const float* src = arena.data() + d.arena;
            for (int i = 0; i < d.boxn[0]; i++)
                for (int j = 0; j < d.boxn[1]; j++)
                    for (int k = 0; k < d.boxn[2]; k++) {
                        const size_t X = (size_t)(d.box0[0] + i), Y = (size_t)(d.box0[1] + j), Z = (size_t)(d.box0[2] + k);
                        col[X + p.bbox_size[0] * (Y + p.bbox_size[1] * Z)] =
                            src[((size_t)i * d.boxn[1] + j) * d.boxn[2] + k];
                    }
        }
        done += R.n;
    }
    return CCCS_OK;
}

extern "C" int cccs_beamlet_dose(cccs_ctx* ctx, const cccs_beam* beams, int nbeams, const float* const* fluence,
                                 float threshold, int mode, cccs_sparse_sink sink, void* user) {
    clear_error();
    if (!ctx || (!beams && nbeams > 0) || nbeams < 0 || (!fluence && nbeams > 0))
        return set_error(CCCS_EINVAL, "null argument");
    if (!sink && mode == CCCS_BEAMLET_PLAIN) return set_error(CCCS_EINVAL, "the plain path needs a sink");
    if (ctx && ctx->nroi > 0 && mode == CCCS_BEAMLET_PLAIN)
        return set_error(CCCS_EINVAL, "ROI reduction runs in the batched modes");
    if (mode != CCCS_BEAMLET_FAITHFUL && mode != CCCS_BEAMLET_PLAIN)
        return set_error(CCCS_EINVAL, "unknown beamlet mode");
    if (!(threshold >= 0.f)) return set_error(CCCS_EINVAL, "threshold must be >= 0");
    CCCS_HIP(hipSetDevice(ctx->device));
    const cccs_problem& p = ctx->prob;
    std::vector<BeamPlan> plans;
    plans.reserve(nbeams);
    for (int i = 0; i < nbeams; i++) {
        if (!fluence[i]) return set_error(CCCS_EINVAL, "null fluence map");
        CCCS_TRY(check_beam(ctx, beams[i]));
        BeamPlan P;
        CCCS_TRY(plan_beam(ctx, beams[i], fluence[i], i, P));
        if (!P.bl.empty()) plans.push_back(std::move(P));
    }
    if (plans.empty()) return CCCS_OK;
    CCCS_TRY(ctx->h_small.ensure(64));
    std::vector<const float*> d_fmaps;
    CCCS_TRY(upload_fmaps(ctx, plans, fluence, d_fmaps));
    const std::vector<BatchRef> refs = all_batches(plans);
    if (mode == CCCS_BEAMLET_PLAIN) {
        for (const BatchRef& R : refs) CCCS_TRY(run_plain_batch(ctx, plans, R, d_fmaps, threshold, sink, user));
        return CCCS_OK;
    }
    const size_t nvox = (size_t)p.bbox_size[0] * p.bbox_size[1] * p.bbox_size[2];
    const size_t budget = size_t(ctx->sched.bl_job_gb) << 30;  // device bytes per job
    const int nset = std::min(std::max(ctx->sched.bl_streams, 1), kRaySets);
    size_t next = 0;
    // device footprint of each batch, additive per cone direction: a job needs its batches' packed
    // grids and sparse-output bounds, plus one direction's REV store, half steps and line table
    // (summed over its batches per direction, maximised over directions)
    const int nrays = ctx->nrays;
    std::vector<std::vector<uint64_t>> b_rev(refs.size()), b_lines(refs.size());
    std::vector<uint64_t> b_fixed(refs.size());
    for (size_t i = 0; i < refs.size(); i++)
        CCCS_TRY(batch_footprint(ctx, plans, refs[i], b_rev[i], b_lines[i], b_fixed[i]));
    // columns of the last finished job, waiting in pinned host memory for the sink
    for (hipEvent_t& e : ctx->stage_ev)
        if (!e) CCCS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (sink && !ctx->cstream) {
        CCCS_HIP(hipStreamCreateWithFlags(&ctx->cstream, hipStreamNonBlocking));
        CCCS_HIP(hipEventCreateWithFlags(&ctx->copy_ev, hipEventDisableTiming));
    }
    hipEvent_t copy_done = nullptr;  // the last job's column copy (its buffers are rewritten after it)
    Delivery pending;
    int njob = 0;
    auto deliver = [&]() -> int {
        if (!pending.live) return CCCS_OK;
        pending.live = false;
        CCCS_HIP(hipEventSynchronize(pending.ev));
        const int nroi = pending.nroi;
        std::vector<uint64_t> blocks((size_t)nroi);
        uint64_t pos = 0;
        for (size_t q = 0; q < pending.cols.size(); q++) {
            const uint32_t* st = pending.stat + 4 * q;
            cccs_sparse_column col{};
            col.beam_index = pending.cols[q].first;
            col.beamlet_index = pending.cols[q].second;
            col.nnz = st[0];
            col.perc_nonzero = (float)(100.0 * (double)st[0] / (double)nvox);
            col.perc_dropped = (float)(100.0 * (double)st[1] / (double)nvox);
            if (nroi > 0) {
                // _reduce_by_roi's statistics (sparse_io.cpp:150-156)
                uint64_t kept = 0;
                for (int r = 0; r < nroi; r++) kept += blocks[(size_t)r] = pending.rcnt[q * (size_t)nroi + (size_t)r];
                const double pn = 100.0 * (double)kept / (double)nvox;
                col.perc_reducedrop = (float)((double)col.perc_nonzero - pn);
                col.perc_nonzero = (float)pn;
                col.nnz = kept;
                col.nblocks = nroi;
                col.row_block_sizes = blocks.data();
            }
            col.keys = pending.keys + pos;
            col.values = pending.vals + pos;
            pos += col.nnz;
            if (sink(user, &col) != 0) return set_error(CCCS_EIO, "sparse sink aborted the call");
        }
        return CCCS_OK;
    };
    while (next < refs.size()) {
        // grow the job batch by batch while it fits the budget
        std::vector<uint64_t> rev(nrays, 0), lines(nrays, 0);
        uint64_t fixed = 0;
        size_t k = next;
        while (k < refs.size()) {
            uint64_t mr = 0, ml = 0;
            for (int r = 0; r < nrays; r++) {
                mr = std::max(mr, rev[r] + b_rev[k][r]);
                ml = std::max(ml, lines[r] + b_lines[k][r]);
            }
            // every direction in flight holds its own REV store, half steps and line table
            const uint64_t bytes =
                fixed + b_fixed[k] + (uint64_t)nset * (sizeof(float) * 2 * mr + sizeof(LineMeta) * ml);
            if (k > next && bytes > budget) break;
            for (int r = 0; r < nrays; r++) {
                rev[r] += b_rev[k][r];
                lines[r] += b_lines[k][r];
            }
            fixed += b_fixed[k];
            k++;
        }
        std::vector<BatchRef> sel(refs.begin() + (long)next, refs.begin() + (long)k);
        Job J;
        CCCS_TRY(build_job(ctx, plans, sel, d_fmaps, J));
        next = k;
        // the previous job's columns reach the sink while this job's cone passes run on the device
        CCCS_TRY(run_job_cones(ctx, J, deliver));
        const int nbl = (int)J.beamlets.size();
        CCCS_TRY(ctx->bl_cnt.ensure(sizeof(uint32_t) * 4 * nbl + sizeof(uint64_t) * (nbl + 1) + 256));
        uint32_t* stat = ctx->bl_cnt.as<uint32_t>();
        uint64_t* offs = reinterpret_cast<uint64_t*>(ctx->bl_cnt.as<uint8_t>() + ((sizeof(uint32_t) * 4 * nbl + 255) & ~size_t(255)));
        const size_t sz_b = sizeof(BatchDev) * J.batches.size();
        ColArgs ca;
        ca.batches = ctx->bl_tab.as<BatchDev>();
        ca.beamlets = reinterpret_cast<const BeamletDev*>(ctx->bl_tab.as<uint8_t>() + ((sz_b + 255) & ~size_t(255)));
        ca.pdose = ctx->bl_pdose.as<float>();
        ca.start = ld3(p.start);
        ca.voxel = ld3(p.voxel);
        ca.rvs = mk3(p.rev_latspacing, p.rev_longspacing, p.rev_latspacing);
        for (int k = 0; k < 3; k++) ca.bbs[k] = (int)p.bbox_start[k];
        ca.nx = (int)p.size[0];
        ca.ny = (int)p.size[1];
        ca.thresh = threshold;
        ca.stat = stat;
        CCCS_TRY(ctx->bl_rowcnt.ensure(sizeof(uint16_t) * std::max<uint64_t>(1, J.rows)));
        ca.rowcnt = ctx->bl_rowcnt.as<uint16_t>();
        ca.off = offs;
        ca.keys = nullptr;
        ca.vals = nullptr;
        if (copy_done) CCCS_HIP(hipStreamWaitEvent(ctx->stream, copy_done, 0));
        int tok = prof_begin(ctx, K_BL_UNPACK);
        hipLaunchKernelGGL(bl_colscan_kernel, dim3(nbl), dim3(256), 0, ctx->stream, ca);
        prof_end(ctx, tok, 0.0, 0.0, (uint64_t)nbl);
        hipLaunchKernelGGL(bl_offsets_kernel, dim3(1), dim3(1024), 0, ctx->stream, stat, nbl, offs);
        CCCS_HIP(hipGetLastError());
        // totals decide the output size
        uint64_t nnz_total = 0;
        CCCS_HIP(hipMemcpyAsync(ctx->h_small.p, offs + nbl, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
        CCCS_HIP(hipStreamSynchronize(ctx->stream));
        std::memcpy(&nnz_total, ctx->h_small.p, sizeof(uint64_t));
        CCCS_TRY(ctx->bl_keys.ensure(sizeof(uint64_t) * std::max<uint64_t>(1, nnz_total)));
        CCCS_TRY(ctx->bl_vals.ensure(sizeof(float) * std::max<uint64_t>(1, nnz_total)));
        ca.keys = ctx->bl_keys.as<uint64_t>();
        ca.vals = ctx->bl_vals.as<float>();
        int tk = prof_begin(ctx, K_BL_COMPACT);
        hipLaunchKernelGGL(bl_emit_kernel, dim3(nbl), dim3(256), 0, ctx->stream, ca);
        // algorithmic traffic: the kept entries (8 B key + 4 B value)
        prof_end(ctx, tk, 12.0 * (double)nnz_total, 0.0, nnz_total);
        CCCS_HIP(hipGetLastError());
        // ROI reduction of the job's columns (count, scan, order-preserving emit)
        const int nroi = ctx->nroi;
        uint64_t rtotal = 0;
        uint32_t* rcnt = nullptr;
        if (nroi > 0) {
            const size_t ncr = (size_t)nbl * (size_t)nroi;
            const size_t cbytes = (sizeof(uint32_t) * ncr + 255) & ~size_t(255);
            CCCS_TRY(ctx->roi_cnt.ensure(cbytes + sizeof(uint64_t) * (ncr + 1)));
            rcnt = ctx->roi_cnt.as<uint32_t>();
            uint64_t* roff = reinterpret_cast<uint64_t*>(ctx->roi_cnt.as<uint8_t>() + cbytes);
            RoiArgs ra;
            ra.rois = ctx->roi_tab.as<RoiDev>();
            ra.rank = ctx->roi_rank.as<int32_t>();
            ra.nroi = nroi;
            ra.nx = (int)p.size[0];
            ra.ny = (int)p.size[1];
            ra.keys = ctx->bl_keys.as<uint64_t>();
            ra.vals = ctx->bl_vals.as<float>();
            ra.off = offs;
            ra.cnt = rcnt;
            ra.roff = roff;
            ra.rkeys = nullptr;
            ra.rvals = nullptr;
            hipLaunchKernelGGL(roi_count_kernel, dim3(nbl), dim3(256), 0, ctx->stream, ra);
            hipLaunchKernelGGL(roi_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, (const uint32_t*)rcnt, ncr, roff);
            CCCS_HIP(hipGetLastError());
            CCCS_HIP(hipMemcpyAsync(ctx->h_small.p, roff + ncr, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
            CCCS_HIP(hipStreamSynchronize(ctx->stream));
            std::memcpy(&rtotal, ctx->h_small.p, sizeof(uint64_t));
            CCCS_TRY(ctx->roi_keys.ensure(sizeof(uint64_t) * std::max<uint64_t>(1, rtotal)));
            CCCS_TRY(ctx->roi_vals.ensure(sizeof(float) * std::max<uint64_t>(1, rtotal)));
            ra.rkeys = ctx->roi_keys.as<uint64_t>();
            ra.rvals = ctx->roi_vals.as<float>();
            hipLaunchKernelGGL(roi_emit_kernel, dim3(nbl), dim3(256), 0, ctx->stream, ra);
            CCCS_HIP(hipGetLastError());
        }
        if (!sink) continue;  // columns stay on the device (overwritten by the next job)
        // columns to pinned host memory (full PCIe rate, one of two buffers in turn); the sink sees
        // them in place, during the next job or at the end of the call
        const uint64_t nout = nroi > 0 ? rtotal : nnz_total;
        const size_t kbytes = sizeof(uint64_t) * nout, vbytes = sizeof(float) * nout;
        const size_t sbytes = sizeof(uint32_t) * 4 * (size_t)nbl;
        const size_t rbytes = sizeof(uint32_t) * (size_t)nbl * (size_t)nroi;
        HostBuf& hbuf = (njob & 1) ? ctx->h_bl_out2 : ctx->h_bl_out;
        CCCS_TRY(hbuf.ensure(kbytes + vbytes + sbytes + rbytes + 64));
        uint8_t* hb = hbuf.as<uint8_t>();
        Delivery& D = pending;
        D.keys = reinterpret_cast<uint64_t*>(hb);
        D.vals = reinterpret_cast<float*>(hb + kbytes);
        D.stat = reinterpret_cast<uint32_t*>(hb + kbytes + vbytes);
        D.rcnt = D.stat + 4 * (size_t)nbl;
        D.nroi = nroi;
        D.cols.resize((size_t)nbl);
        for (int q = 0; q < nbl; q++) {
            const BeamletDev& d = J.beamlets[q];
            D.cols[(size_t)q] = {plans[sel[d.batch].plan].beam, d.fidx};
        }
        // the copies run on their own stream: the next job's cone passes, queued behind this
        // compaction, do not wait for them; the next compaction waits before rewriting the buffers
        hipStream_t cs = ctx->cstream;
        CCCS_HIP(hipEventRecord(ctx->copy_ev, ctx->stream));
        CCCS_HIP(hipStreamWaitEvent(cs, ctx->copy_ev, 0));
        CCCS_HIP(hipMemcpyAsync(D.stat, stat, sbytes, hipMemcpyDeviceToHost, cs));
        if (nroi > 0) CCCS_HIP(hipMemcpyAsync(D.rcnt, rcnt, rbytes, hipMemcpyDeviceToHost, cs));
        if (nout) {
            CCCS_HIP(hipMemcpyAsync(D.keys, nroi > 0 ? ctx->roi_keys.p : ctx->bl_keys.p, kbytes, hipMemcpyDeviceToHost,
                                    cs));
            CCCS_HIP(hipMemcpyAsync(D.vals, nroi > 0 ? ctx->roi_vals.p : ctx->bl_vals.p, vbytes, hipMemcpyDeviceToHost,
                                    cs));
        }
        CCCS_HIP(hipEventRecord(ctx->stage_ev[njob & 1], cs));
        D.ev = ctx->stage_ev[njob & 1];
        copy_done = D.ev;
        D.live = true;
        njob++;
    }
    CCCS_TRY(deliver());
    return CCCS_OK;
}
