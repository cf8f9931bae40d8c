// walk_check.hip -- CPU check of the cone kernel's segment-sum walks.  TEST INFRASTRUCTURE.
//
// Builds random REV line segments on the host (water, lung, bone and air runs; TERMA with zero
// runs), sets up the segment-sum tables exactly as the cone kernel lays them out in LDS (FP32
// radiological centres, FP64 prefix sums E/E2, the lookup table built by the kernel's per-sample
// scatter, checked entry by entry against its definition), and compares for every destination
// sample the segment-sum walks (ramp_back + ramp_fwd) with the sequential per-sample walk
// (walk_both, the reference's RowConvolve loop): within 2e-3 of the line's largest value (the
// sequential walk accumulates the radiological distance in FP32, which drifts over long air runs
// of tiny steps; the segment-sum walk's FP64 prefix sums do not).
// This file includes endtoend_amd/csrc/beam.hip and calls its functions on the host.
//   walk_check [lines] [seed]
// Prints one JSON line of counts and exits 1 on any mismatch.
#include "../../endtoend_amd/csrc/beam.hip"

#include <cinttypes>
#include <random>

namespace {

constexpr float kRadii[24] = {0.05f, 0.1f, 0.15f, 0.2f, 0.3f, 0.4f, 0.5f, 0.6f, 0.8f, 1.0f, 1.5f, 2.0f,
                              3.0f,  4.0f, 5.0f,  6.0f, 8.0f, 10.f, 15.f, 20.f, 30.f, 40.f, 50.f, 60.f};

struct Line {
    std::vector<float2> s;
    std::vector<float> c;
    std::vector<double> E, E2;
    std::vector<uint16_t> lb;
    RampLine L;
    uint64_t table_errors = 0;
};

// The cone kernel's LDS set-up of a fully valid segment, done sequentially.
void setup(Line& ln) {
    const int n = (int)ln.s.size();
    ln.c.assign(n + 4, 0.f);
    ln.E.assign(n + 1, 0.0);
    ln.E2.assign(n + 1, 0.0);
    double P = 0.0;
    for (int k = 0; k < n; k++) {
        ln.c[k] = (float)(P + (double)ln.s[k].x);
        P += 2.0 * (double)ln.s[k].x;
    }
    for (int q = 0; q < 4; q++) ln.c[n + q] = INFINITY;
    double e1 = 0.0, e2 = 0.0;
    for (int k = 0; k < n; k++) {
        const double ck = (double)ln.c[k], t = (double)ln.s[k].y;
        const double tp = k > 0 ? (double)ln.s[k - 1].y : 0.0, tn = k + 1 < n ? (double)ln.s[k + 1].y : 0.0;
        ln.E[k] = e1;
        ln.E2[k] = e2;
        e1 += (t - tp) * ck;
        e2 += (t - tn) * ck;
    }
    ln.E[n] = e1;
    ln.E2[n] = e2;
    RampLine& L = ln.L;
    L.c = ln.c.data();
    L.E = ln.E.data();
    L.E2 = ln.E2.data();
    L.s = ln.s.data();
    L.n = n;
    L.U = 2 * n;
    L.c0 = ln.c[0];
    L.cn = ln.c[n - 1];
    const float range = L.cn - L.c0;
    const float D = range > 0.f ? range / (float)L.U : 0.f;
    L.invD = range > 0.f ? (float)L.U / range : 0.f;
    // the kernel's per-sample scatter, checked against the table's definition (binary search)
    ln.lb.assign(L.U + 2, 0xffff);
    for (int k = 0; k <= n; k++) {
        const int u0 = k == 0 ? 0 : ramp_cell_after(L.c0, D, L.invD, L.U + 2, ln.c[k - 1]);
        const int u1 = k == n ? L.U + 2 : ramp_cell_after(L.c0, D, L.invD, L.U + 2, ln.c[k]);
        for (int u = u0; u < u1; u++) {
            if (ln.lb[u] != 0xffff) ln.table_errors++;
            ln.lb[u] = (uint16_t)k;
        }
    }
    for (int u = 0; u <= L.U + 1; u++) {
        const float t = L.c0 + (float)u * D;
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ln.c[mid] < t) lo = mid + 1; else hi = mid;
        }
        if (ln.lb[u] != (uint16_t)lo) ln.table_errors++;
    }
    L.lb = ln.lb.data();
}

}  // namespace

int main(int argc, char** argv) {
    const int nlines = argc > 1 ? std::atoi(argv[1]) : 2000;
    const unsigned seed = argc > 2 ? (unsigned)std::atoi(argv[2]) : 7u;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U01(0.f, 1.f);
    const int nr = 24;
    const float delr = 0.2f;
    // two cumulative kernel rows (forward, backward) rising to ~1 over the radii
    float rowF[nr], rowB[nr], IW[nr];
    for (int q = 0; q < nr; q++) {
        rowF[q] = 1.f - std::exp(-kRadii[q] / 1.7f) * (0.6f + 0.4f * std::exp(-kRadii[q] / 0.3f));
        rowB[q] = 0.4f * (1.f - std::exp(-kRadii[q] / 0.9f));
        IW[q] = 1.f / (kRadii[q] - (q == 0 ? 0.f : kRadii[q - 1]));
    }
    // the kernel rows as C(rho) = beta_s + sigma_s rho, formed as the cone kernel forms them
    double sig[2][nr], beta[2][nr];
    for (int d = 0; d < 2; d++)
        for (int q = 0; q < nr; q++) {
            const float* row = d == 0 ? rowF : rowB;
            const float prev = q == 0 ? 0.f : row[q - 1];
            const float sg = (row[q] - prev) * IW[q];
            sig[d][q] = (double)sg;
            beta[d][q] = (double)prev - (double)sg * (q == 0 ? 0.0 : (double)kRadii[q - 1]);
        }
    const float dens[4] = {1.0f, 0.25f, 1.85f, 0.001f};
    uint64_t dests = 0, walk_mismatch = 0, table_errors = 0, iters = 0;
    double worst = 0.0;
    std::string first;
    for (int li = 0; li < nlines; li++) {
        Line ln;
        const int n = 1 + (int)(U01(rng) * 400.f);
        ln.s.resize(n);
        int k = 0;
        float t = 0.5f + U01(rng);
        while (k < n) {  // material runs
            const int run = 1 + (int)(U01(rng) * 60.f);
            const int m = (int)(U01(rng) * 4.f) & 3;
            const bool zero = U01(rng) < 0.15f;
            for (int q = 0; q < run && k < n; q++, k++) {
                t *= 0.97f + 0.02f * U01(rng);
                ln.s[k] = make_float2(0.5f * (delr * dens[m]), zero ? 0.f : t);
            }
        }
        setup(ln);
        table_errors += ln.table_errors;
        std::vector<double> v(n);
        std::vector<float> vw(n);
        double vmax = 0.0;
        for (int j = 0; j < n; j++) {
            uint32_t a = 0, w = 0;
            v[j] = ramp_back(ln.L, j, sig[0], beta[0], kRadii, nr, a) + ramp_fwd(ln.L, j, sig[1], beta[1], kRadii, nr, a);
            vw[j] = walk_both(ln.s.data(), j, 0, n - 1, kRadii, IW, rowF, rowB, nr, true, w);
            iters += a;
            vmax = std::max(vmax, std::fabs(v[j]));
        }
        for (int j = 0; j < n; j++) {
            const double d = std::fabs((double)vw[j] - v[j]) / std::max(vmax, 1e-30);
            worst = std::max(worst, d);
            if (d > 2e-3) {
                if (!walk_mismatch) {
                    char buf[256];
                    std::snprintf(buf, sizeof(buf), "line %d n %d j %d: %.9g vs %.9g", li, n, j, v[j], (double)vw[j]);
                    first = buf;
                }
                walk_mismatch++;
            }
        }
        dests += (uint64_t)n;
    }
    std::printf("{\"lines\": %d, \"destinations\": %" PRIu64 ", \"walk_mismatch\": %" PRIu64
                ", \"worst_walk_rel\": %.3e, \"table_errors\": %" PRIu64 ", \"iters\": %" PRIu64
                ", \"first\": \"%s\"}\n",
                nlines, dests, walk_mismatch, worst, table_errors, iters, first.c_str());
    return walk_mismatch || table_errors ? 1 : 0;
}
