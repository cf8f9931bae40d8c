// beam.hip -- resident context and the plain full-beam CCCS path on gfx950.
//
// Per beam (replacing radconvolveTexture, CCCS/dosecalc-beam/nvbbRayConvolve_host.cu:376-725):
//   1. terma_kernel      Siddon radiological path + poly-energetic TERMA per calc_bbox voxel
//                        (DoseCalcAlgo/cudaSiddon.cu:117-308)
//   2. crop_partial/final  BEV-rotated extent of non-zero TERMA, reduced on the GPU instead of the
//                        reference's D2H copy + host triple loop (:417-499); 28 bytes come back
//   3. host geometry     crop box + one REV grid per cone (findREV :75-155)
//   4. cone_kernel       one workgroup per REV line, all cones in one launch: resample density and
//                        TERMA along the line into LDS (cudaRaytrace, nvbbRayConvolve_device.cu:11-80)
//                        then the collapsed-cone walk in both directions from LDS (RowConvolve
//                        :87-218).  The REV grid is sized per cone; nothing is cleared per cone.
//   5. pvcs_kernel       one pass over the PVCS voxels the crop box can reach, gathering every
//                        cone's REV dose trilinearly and adding the sum once (REV2XYZdose
//                        :222-261 per cone in the reference).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <array>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "context.hpp"
#include "device_common.hpp"

namespace cccs {

BeamDev make_beam_dev(const cccs_beam& b) {
    BeamDev d;
    d.src = ld3(b.source);
    d.dir_n = norm3(ld3(b.direction));
    d.iso = ld3(b.isocenter);
    d.sad = b.sad;
    d.iso_dist = len3(d.iso - d.src);
    d.fmx = (int)b.fmap_size[0];
    d.fmy = (int)b.fmap_size[1];
    d.bsx = b.beamlet_size[0];
    d.bsy = b.beamlet_size[1];
    d.beam_fwd = rot2_beam_fwd(b.azimuth, b.zenith, b.coll);
    return d;
}

int check_beam(const cccs_ctx* ctx, const cccs_beam& b) {
    (void)ctx;
    if (b.fmap_size[0] == 0 || b.fmap_size[1] == 0 || b.fmap_size[0] > 4096 || b.fmap_size[1] > 4096)
        return set_error(CCCS_EINVAL, "fluence map dims must be in [1, 4096]");
    if (!(b.beamlet_size[0] > 0.f) || !(b.beamlet_size[1] > 0.f)) return set_error(CCCS_EINVAL, "beamlet size must be > 0");
    if (!(b.sad > 0.f)) return set_error(CCCS_EINVAL, "sad must be > 0");
    return CCCS_OK;
}

// ------------------------------------------------------------------------------------------------
// kernels

struct TermaArgs {
    float* terma;
    const float* fmap;
    const float* dens;
    VolDesc vd;
    BeamDev beam;
    SpecDev spec;
    int bx, by, bz;          // calc_bbox size
    int ox, oy, oz;          // calc_bbox start
    int ss;                  // supersampling factor
    float bhc;
    int beamletnum;
    uint32_t* counts;         // per-block Siddon step counts (profiling) or nullptr
};

// Centre of calc_bbox voxel (X, Y, Z).
CCCS_HD v3 terma_point(const TermaArgs& a, int X, int Y, int Z) {
    return mk3(a.vd.start.x + ((float)a.ox + X) * a.vd.voxel.x, a.vd.start.y + ((float)a.oy + Y) * a.vd.voxel.y,
               a.vd.start.z + ((float)a.oz + Z) * a.vd.voxel.z);
}

// One lane per calc_bbox voxel (x fastest => coalesced store).  Each lane marches its own ray
// (the rpl-only march: same crossings and segments as cudaSiddon, fetch coordinates formed with
// reciprocal voxel sizes); neighbouring lanes follow near-parallel rays, so the density taps stay
// in L1/L2.
__global__ void __launch_bounds__(256) terma_kernel(TermaArgs a) {
    const int X = blockIdx.x * 64 + threadIdx.x;
    const int Y = blockIdx.y * 4 + threadIdx.y;
    const int Z = blockIdx.z;
    const bool in = (X < a.bx && Y < a.by && Z < a.bz);
    uint32_t steps = 0;
    float acc = 0.f;
    const v3 ivox = mk3(1.f / a.vd.voxel.x, 1.f / a.vd.voxel.y, 1.f / a.vd.voxel.z);
    if (!in) {
    } else if (a.ss > 1) {
        // supersampled: sub-voxel centres, summed in sub-voxel order then averaged
        const int nsub = a.ss * a.ss * a.ss;
        v3 subsz = mk3(a.vd.voxel.x / (float)a.ss, a.vd.voxel.y / (float)a.ss, a.vd.voxel.z / (float)a.ss);
        acc = 0.f;
        for (int s = 0; s < nsub; s++) {
            int sx = s % a.ss, sy = (s / a.ss) % a.ss, sz = s / (a.ss * a.ss);
            v3 p = mk3(a.vd.start.x + ((float)a.ox + X - 0.5f) * a.vd.voxel.x + (sx + 0.5f) * subsz.x,
                       a.vd.start.y + ((float)a.oy + Y - 0.5f) * a.vd.voxel.y + (sy + 0.5f) * subsz.y,
                       a.vd.start.z + ((float)a.oz + Z - 0.5f) * a.vd.voxel.z + (sz + 0.5f) * subsz.z);
            int it = 0;
            acc += point_terma_rpl(p, a.fmap, a.beam, a.spec, a.vd, ivox, a.dens, a.bhc, a.beamletnum, &it);
            steps += (uint32_t)it;
        }
        acc /= (float)nsub;
    } else {
        // fluence times the unit-fluence TERMA (the Siddon march only inside the field), the order
        // in which a resident unit-TERMA grid of the same source gives the same value
        const v3 p = terma_point(a, X, Y, Z);
        const float fl = point_fluence(p, a.fmap, a.beam, a.beamletnum);
        int it = 0;
        acc = fl > 0.f ? fl * unit_point_terma(p, a.beam, a.spec, a.vd, ivox, a.dens, a.bhc, &it) : 0.f;
        steps = (uint32_t)it;
    }
    if (in) a.terma[(size_t)X + (size_t)a.bx * ((size_t)Y + (size_t)a.by * Z)] = acc;
    if (a.counts) block_count(steps, a.counts, blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z));
}

// Unit-fluence TERMA of every calc_bbox voxel (the march runs for every voxel): one beam source's
// entry of the angle cache.
__global__ void __launch_bounds__(256) unit_terma_kernel(TermaArgs a, float* __restrict__ unit) {
    const int X = blockIdx.x * 64 + threadIdx.x;
    const int Y = blockIdx.y * 4 + threadIdx.y;
    const int Z = blockIdx.z;
    if (X >= a.bx || Y >= a.by || Z >= a.bz) return;
    const v3 ivox = mk3(1.f / a.vd.voxel.x, 1.f / a.vd.voxel.y, 1.f / a.vd.voxel.z);
    unit[(size_t)X + (size_t)a.bx * ((size_t)Y + (size_t)a.by * Z)] =
        unit_point_terma(terma_point(a, X, Y, Z), a.beam, a.spec, a.vd, ivox, a.dens, a.bhc, nullptr);
}

// TERMA from a resident unit-TERMA grid of the beam's source: the voxel's fluence times its entry.
__global__ void __launch_bounds__(256) terma_from_unit_kernel(TermaArgs a, const float* __restrict__ unit) {
    const int X = blockIdx.x * 64 + threadIdx.x;
    const int Y = blockIdx.y * 4 + threadIdx.y;
    const int Z = blockIdx.z;
    if (X >= a.bx || Y >= a.by || Z >= a.bz) return;
    const size_t v = (size_t)X + (size_t)a.bx * ((size_t)Y + (size_t)a.by * Z);
    const float fl = point_fluence(terma_point(a, X, Y, Z), a.fmap, a.beam, a.beamletnum);
    a.terma[v] = fl > 0.f ? fl * unit[v] : 0.f;
}

struct PathArgs {
    const float* dens;
    VolDesc vd;
    v3 src;
    const float* pts;
    int n, maxseg;
    int32_t* ijk;
    float* mid;
    float* len;
    int32_t* counts;
    float* rpl;
};

__global__ void __launch_bounds__(256) siddon_path_kernel(PathArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    int c = 0;
    v3 p = mk3(a.pts[3 * i], a.pts[3 * i + 1], a.pts[3 * i + 2]);
    float r = siddon_march(p, a.vd.start, a.vd.end, a.vd.voxel, a.src, a.dens, a.vd.nx, a.vd.ny, a.vd.nz,
                           [&](float fx, float fy, float fz, float l) {
                               if (c < a.maxseg) {
                                   size_t o = (size_t)i * a.maxseg + c;
                                   a.ijk[3 * o] = f2i_rd(fx + 0.5f);
                                   a.ijk[3 * o + 1] = f2i_rd(fy + 0.5f);
                                   a.ijk[3 * o + 2] = f2i_rd(fz + 0.5f);
                                   a.mid[3 * o] = fx;
                                   a.mid[3 * o + 1] = fy;
                                   a.mid[3 * o + 2] = fz;
                                   a.len[o] = l;
                               }
                               c++;
                           },
                           nullptr);
    a.counts[i] = c;
    a.rpl[i] = r;
}

// Crop box: min/max of the BEV-rotated bbox index of every voxel with non-zero TERMA.
// Pass 1 writes one partial (min xyz, max xyz, count) per workgroup.
constexpr int kCropBlocks = 1024;
__global__ void __launch_bounds__(256) crop_partial_kernel(const float* __restrict__ terma, int bx, int by, int bz,
                                                           Rot2 rot, float* __restrict__ part) {
    __shared__ float s[7][256];
    float mn0 = FLT_MAX, mn1 = FLT_MAX, mn2 = FLT_MAX, mx0 = -FLT_MAX, mx1 = -FLT_MAX, mx2 = -FLT_MAX, cnt = 0.f;
    const size_t n = (size_t)bx * by * bz;
    for (size_t v = (size_t)blockIdx.x * 256 + threadIdx.x; v < n; v += (size_t)gridDim.x * 256) {
        if (terma[v] != 0.f) {
            int x = (int)(v % bx);
            int y = (int)((v / bx) % by);
            int z = (int)(v / ((size_t)bx * by));
            v3 b = rot2(rot, mk3((float)x, (float)y, (float)z));
            mn0 = fminf(mn0, b.x); mn1 = fminf(mn1, b.y); mn2 = fminf(mn2, b.z);
            mx0 = fmaxf(mx0, b.x); mx1 = fmaxf(mx1, b.y); mx2 = fmaxf(mx2, b.z);
            cnt += 1.f;
        }
    }
    const int t = threadIdx.x;
    s[0][t] = mn0; s[1][t] = mn1; s[2][t] = mn2; s[3][t] = mx0; s[4][t] = mx1; s[5][t] = mx2; s[6][t] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            for (int q = 0; q < 3; q++) s[q][t] = fminf(s[q][t], s[q][t + w]);
            for (int q = 3; q < 6; q++) s[q][t] = fmaxf(s[q][t], s[q][t + w]);
            s[6][t] += s[6][t + w];
        }
        __syncthreads();
    }
    if (t < 7) part[(size_t)blockIdx.x * 8 + t] = s[t][0];
}

__global__ void __launch_bounds__(256) crop_final_kernel(const float* __restrict__ part, int nblocks, float* __restrict__ out) {
    __shared__ float s[7][256];
    const int t = threadIdx.x;
    float v[7] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX, 0.f};
    for (int b = t; b < nblocks; b += 256) {
        for (int q = 0; q < 3; q++) v[q] = fminf(v[q], part[(size_t)b * 8 + q]);
        for (int q = 3; q < 6; q++) v[q] = fmaxf(v[q], part[(size_t)b * 8 + q]);
        v[6] += part[(size_t)b * 8 + 6];
    }
    for (int q = 0; q < 7; q++) s[q][t] = v[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            for (int q = 0; q < 3; q++) s[q][t] = fminf(s[q][t], s[q][t + w]);
            for (int q = 3; q < 6; q++) s[q][t] = fmaxf(s[q][t], s[q][t + w]);
            s[6][t] += s[6][t + w];
        }
        __syncthreads();
    }
    if (t < 7) out[t] = s[t][0];
}

// A class-list entry: the line, its cone and its record, so that a workgroup starts from one
// 16-byte load instead of a list load, a record load and a search over the cones.
struct ConeLine {
    uint32_t line;
    uint32_t cone;
    LineMeta meta;
};

struct ConeArgs {
    const float2* dt;         // {density, TERMA (0 outside the calc_bbox)} over the CT grid
    const float* tables;      // kernel | radii | inverse widths
    const ConeDev* cones;
    LineMeta* lines;          // one entry per REV line of every cone
    float* rev;               // compact REV dose store
    uint64_t nlines;
    int ncones;
    int ntheta, nradii, nphi;
    float delr;
    uint32_t* counts;         // per-workgroup walk iteration counts (profiling) or nullptr
    int ramp_min;             // valid segments at least this long use the segment-sum walk
    int maxn;                 // longest stored line of this launch (LDS sizing)
    int len_lo, len_hi;       // this launch takes the lines whose stored length is in [len_lo, len_hi]
    const ConeLine* list;     // the lines of this launch (one workgroup each), or nullptr: every line
    VolDesc vd;
    m33 binv;                 // PVCS <- BEV rotation (index units)
    v3 off;                   // (bevStart - start) / voxel
    v3 bev_size;              // crop-box size (index units)
    v3 bb0, bb1;              // calc_bbox bounds (PVCS index)
    int bbx, bby, bbz;        // calc_bbox size
};

__device__ __forceinline__ int cone_of_line(const ConeDev* __restrict__ cones, int ncones, uint64_t line) {
    int lo = 0, hi = ncones - 1;
    while (lo < hi) {  // last cone whose first line <= line
        int mid = (lo + hi + 1) >> 1;
        if (cones[mid].line0 <= line) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Sample interval [i0, i1) of REV line (tY, tZ) of cone c that can lie inside the crop box and the
// calc_bbox: the BEV and PVCS positions along a line are affine in the sample index, so each box
// clips the line to a slab interval.  Widened by 2/3 samples; the exact per-sample test of
// cudaRaytrace (nvbbRayConvolve_device.cu:50-70) still decides membership.
__device__ __forceinline__ void line_interval(const ConeArgs& a, const ConeDev& c, int tY, int tZ, int& i0,
                                              int& i1) {
    const int n = c.nx;
    const v3 p0 = mk3(c.rs.x + (float)tZ, c.rs.y, c.rs.z + (float)tY);
    const v3 b0 = c.kinv(p0);
    const v3 kd = mk3(c.kinv.m[1], c.kinv.m[4], c.kinv.m[7]);
    const v3 g0 = a.binv(b0) + a.off;
    const v3 gd = a.binv(kd);
    float t0 = 0.f, t1 = (float)n;
    clip_slab(b0.x, kd.x, 0.f, a.bev_size.x, t0, t1);
    clip_slab(b0.y, kd.y, 0.f, a.bev_size.y, t0, t1);
    clip_slab(b0.z, kd.z, 0.f, a.bev_size.z, t0, t1);
    clip_slab(g0.x, gd.x, a.bb0.x, a.bb1.x, t0, t1);
    clip_slab(g0.y, gd.y, a.bb0.y, a.bb1.y, t0, t1);
    clip_slab(g0.z, gd.z, a.bb0.z, a.bb1.z, t0, t1);
    i0 = 0;
    i1 = 0;
    if (t0 <= t1) {
        i0 = max(0, (int)floorf(t0) - 2);
        i1 = min(n, (int)ceilf(t1) + 3);
        if (i1 < i0) i1 = i0;
    }
}

// ---- segment-sum form of the two walks --------------------------------------------------------
// walk_both visits every sample of the segment for every destination (O(line^2) per line).  The
// same sums follow in O(nradii) per destination from three facts:
//  * the radiological distance from destination j to sample k is a difference of the line's
//    radiological coordinates: rho_k = P - c_k backward (P = boundary between j-1 and j, c_k the
//    centre of sample k) and c_k - P forward (P = boundary between j and j+1);
//  * the cumulative kernel C is linear on each radial interval (R_{s-1}, R_s]:
//    C(rho) = beta_s + sigma_s * rho;
//  * summing by parts, sum_k T_k (C(rho_k) - C(rho_{k+1})) = sum_k (T_k - T_{k-1}) C(rho_k) + one
//    end term, so a run of samples in one interval contributes
//    (beta_s + sigma_s P)(T_hi - T_{lo-1}) - sigma_s (E[hi+1] - E[lo]) with E the prefix sum of
//    (T_k - T_{k-1}) c_k (FP64).
// The samples of each interval are found by a lookup in a uniform table over the radiological
// coordinate.  The walk's exits (distance beyond the last radius, entering the last interval) become
// the first/last sample of the sum.  Same result as the sequential walk up to FP32 rounding of the
// accumulated distance.
struct RampLine {
    const float* c;       // centres, radiological cm from the segment start; c[n..n+3] = +inf
    const double* E;      // E[k]  = sum_{m<k} (T_m - T_{m-1}) c_m   (T_{-1} = 0), k in [0, n]
    const double* E2;     // E2[k] = sum_{m<k} (T_m - T_{m+1}) c_m   (T_n = 0),    k in [0, n]
    const float2* s;      // {half step, TERMA} of the segment's samples
    const uint16_t* lb;   // lb[u] = first k with c_k >= c0 + u * D, u in [0, U + 1] (D = range / U)
    int n, U;
    float c0, cn, invD;
};

// first k with c_k >= y
CCCS_HD int ramp_lb(const RampLine& L, float y) {
    if (!(y > L.c0)) return 0;
    if (y > L.cn) return L.n;
    // biased by 1% of a cell so that the cell's threshold never exceeds y after rounding
    int u = (int)((y - L.c0) * L.invD - 0.01f);
    u = min(max(u, 0), L.U - 1);
    int k = L.lb[u];
    if (L.lb[u + 1] - k > 8) {  // a crowded cell (low density): bisect up to the next cell's end
        int hi = L.lb[u + 2];
        while (k < hi) {
            const int mid = (k + hi) >> 1;
            if (L.c[mid] < y) k = mid + 1; else hi = mid;
        }
    }
    while (L.c[k] < y) k++;
    return k;
}

// first k with c_k > y
CCCS_HD int ramp_ub(const RampLine& L, float y) {
    if (y < L.c0) return 0;
    if (!(y < L.cn)) return L.n;
    int u = (int)((y - L.c0) * L.invD - 0.01f);
    u = min(max(u, 0), L.U - 1);
    int k = L.lb[u];
    if (L.lb[u + 1] - k > 8) {
        int hi = L.lb[u + 2];
        while (k < hi) {
            const int mid = (k + hi) >> 1;
            if (L.c[mid] <= y) k = mid + 1; else hi = mid;
        }
    }
    while (L.c[k] <= y) k++;
    return k;
}

// Backward walk of destination j (toward lower indices), kernel row given as sigma/beta.
CCCS_HD double ramp_back(const RampLine& L, int j, const double* __restrict__ sig,
                                            const double* __restrict__ beta, const float* __restrict__ R, int nr,
                                            uint32_t& iters) {
    if (j == 0) return 0.0;
    const float2 ep = L.s[j - 1];
    const double P = (double)L.c[j - 1] + (double)ep.x;
    const int nlast = ramp_lb(L, (float)(P - (double)R[nr - 1]));
    const int nprev = ramp_lb(L, (float)(P - (double)R[nr - 2]));
    const int a = max(nlast, nprev - 1);  // last sample the walk reaches
    if (a >= j) return 0.0;
    double sum = 0.0, Thi = (double)ep.y, Ehi = L.E[j];
    int hiend = j;
    for (int s = 0; s < nr; s++) {
        int ns = s == nr - 1 ? nlast : (s == nr - 2 ? nprev : ramp_lb(L, (float)(P - (double)R[s])));
        ns = max(ns, a);
        if (ns < hiend) {  // samples [ns, hiend) lie in interval s
            const double Tlo = ns > 0 ? (double)L.s[ns - 1].y : 0.0;
            const double Elo = L.E[ns];
            sum += (beta[s] + sig[s] * P) * (Thi - Tlo) - sig[s] * (Ehi - Elo);
            if (ns == a) sum += Tlo * (beta[s] + sig[s] * (P - (double)L.c[a]));
            hiend = ns;
            Thi = Tlo;
            Ehi = Elo;
            iters++;
        }
        if (hiend == a) break;
    }
    return sum;
}

// Forward walk of destination j (toward higher indices).
CCCS_HD double ramp_fwd(const RampLine& L, int j, const double* __restrict__ sig,
                                           const double* __restrict__ beta, const float* __restrict__ R, int nr,
                                           uint32_t& iters) {
    if (j >= L.n - 1) return 0.0;
    const float2 en = L.s[j + 1];
    const double P = (double)L.c[j + 1] - (double)en.x;
    const int mlast = ramp_ub(L, (float)(P + (double)R[nr - 1])) - 1;
    const int mprev = ramp_ub(L, (float)(P + (double)R[nr - 2])) - 1;
    const int b = min(mlast, mprev + 1);  // last sample the walk reaches
    if (b <= j) return 0.0;
    double sum = 0.0, Tlo = (double)en.y, Elo = L.E2[j + 1];
    int lo = j + 1;
    for (int s = 0; s < nr; s++) {
        int ms = s == nr - 1 ? mlast : (s == nr - 2 ? mprev : ramp_ub(L, (float)(P + (double)R[s])) - 1);
        ms = min(ms, b);
        if (ms >= lo) {  // samples [lo, ms] lie in interval s
            const double Thi = ms + 1 < L.n ? (double)L.s[ms + 1].y : 0.0;
            const double Ehi = L.E2[ms + 1];
            sum += (beta[s] - sig[s] * P) * (Tlo - Thi) + sig[s] * (Ehi - Elo);
            if (ms == b) sum += Thi * (beta[s] + sig[s] * ((double)L.c[b] - P));
            lo = ms + 1;
            Tlo = Thi;
            Elo = Ehi;
            iters++;
        }
        if (lo > b) break;
    }
    return sum;
}

// First cell u in [0, nu] whose threshold c0 + u * D exceeds v (nu when none does): an estimate
// from the cell width, corrected against the thresholds themselves (as the table defines them).
CCCS_HD int ramp_cell_after(float c0, float D, float invD, int nu, float v) {
    if (!(D > 0.f)) return v < c0 ? 0 : nu;
    const float g = fminf(fmaxf((v - c0) * invD, -1.f), (float)nu);
    int u = min(max((int)floorf(g) + 1, 0), nu);
    while (u > 0 && c0 + (float)(u - 1) * D > v) u--;
    while (u < nu && !(c0 + (float)u * D > v)) u++;
    return u;
}

// Exclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ double wave_excl_sum(double v) {
    const int lane = threadIdx.x & 63;
    double x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const double u = __shfl_up(x, o, 64);
        if (lane >= o) x += u;
    }
    const double e = __shfl_up(x, 1, 64);
    return lane == 0 ? 0.0 : e;
}

// LDS bytes of the cone kernel for lines of up to maxn samples (with the segment-sum arrays).
inline size_t cone_lds_bytes(uint32_t maxn, bool ramp) {
    if (!ramp) return sizeof(float2) * maxn;
    const size_t n1 = (size_t)maxn + 1, nc = ((size_t)maxn + 4 + 1) & ~(size_t)1;
    return 2 * sizeof(double) * n1 + sizeof(float) * nc + sizeof(float2) * maxn + sizeof(uint16_t) * (2 * maxn + 2);
}

// Pass 1 of the cone pipeline: the stored interval of every REV line (offsets come from a scan).
__global__ void __launch_bounds__(256) line_interval_kernel(ConeArgs a) {
    const uint64_t line = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (line >= a.nlines) return;
    const int q = cone_of_line(a.cones, a.ncones, line);
    const ConeDev& c = a.cones[q];
    const uint64_t l = line - c.line0;
    int i0, i1;
    line_interval(a, c, (int)(l % (uint64_t)c.ny), (int)(l / (uint64_t)c.ny), i0, i1);
    LineMeta m;
    m.i0 = (int16_t)i0;
    m.len = (int16_t)(i1 - i0);
    m.off = 0;
    a.lines[line] = m;
}

// Pass 2: one wave per REV line of any cone (XCD-aware order).  The stored interval of the line is
// resampled (density and TERMA, trilinear) into LDS with the exact per-sample test of
// cudaRaytrace; then every valid sample runs both collapsed-cone walks from LDS and the interval's
// cone dose is written to the compact REV store.
// Workgroup-wide exclusive prefix sum of one double per thread (NT/64 waves; s_tot holds NT/64
// doubles).  Contains barriers: every thread of the workgroup must call it.
template <int NT>
__device__ __forceinline__ double block_excl_sum(double v, double* s_tot) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const double u = __shfl_up(x, o, 64);
        if (lane >= o) x += u;
    }
    double e = __shfl_up(x, 1, 64);
    e = lane == 0 ? 0.0 : e;
    if (NT == 64) return e;
    if (lane == 63) s_tot[wid] = x;
    __syncthreads();
    double base = 0.0;
    for (int w = 0; w < wid; w++) base += s_tot[w];
    __syncthreads();
    return base + e;
}

template <int NT>
__global__ void __launch_bounds__(NT) cone_kernel(ConeArgs a) {
    constexpr int NW = NT / 64;
    extern __shared__ double s_dyn[];
    __shared__ float sR[kMaxRadii], sIW[kMaxRadii], sF[kMaxRadii], sB[kMaxRadii];
    __shared__ double sSig[2][kMaxRadii], sBeta[2][kMaxRadii];
    __shared__ double s_tot[NW];
    __shared__ int s_red[NW][3];
    __shared__ uint32_t s_cnt[NW];
    const uint64_t slot = xcd_remap(blockIdx.x, gridDim.x);
    uint64_t line;
    uint32_t cone;
    LineMeta meta;
    if (a.list) {
        const ConeLine e = a.list[slot];
        line = e.line;
        cone = e.cone;
        meta = e.meta;
    } else {
        line = slot;
        meta = a.lines[line];
        cone = (uint32_t)cone_of_line(a.cones, a.ncones, line);
    }
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const bool mine = meta.len >= a.len_lo && meta.len <= a.len_hi;
    uint32_t iters = 0;
    if (mine) {
        const ConeDev& c = a.cones[cone];
        const int nr = a.nradii;
        const int m = meta.len;
        const bool ramp_lds = a.ramp_min <= a.maxn;
        // LDS: [E | E2 | c | line | lb] with the segment-sum arrays, else [line]
        const int n1 = a.maxn + 1, nc = (a.maxn + 5) & ~1;
        double* rE = s_dyn;
        double* rE2 = rE + n1;
        float* rc = reinterpret_cast<float*>(rE2 + n1);
        float2* s_line = ramp_lds ? reinterpret_cast<float2*>(rc + nc) : reinterpret_cast<float2*>(s_dyn);
        uint16_t* rlb = reinterpret_cast<uint16_t*>(s_line + a.maxn);
        // the kernel rows' loads are issued before the resample's and land in LDS after it, so
        // their latency overlaps the density / TERMA fetches
        float tR = 0.f, tIW = 0.f, tF = 0.f, tB = 0.f;
        if (tid < nr) {
            const float* K = a.tables;
            const float* R = a.tables + (size_t)a.ntheta * nr;
            tR = R[tid];
            tIW = R[nr + tid];
            tF = K[(size_t)c.rowf * nr + tid];
            tB = K[(size_t)c.rowb * nr + tid];
        }
        const uint64_t l = line - c.line0;
        const int tY = (int)(l % (uint64_t)c.ny);
        const int tZ = (int)(l / (uint64_t)c.ny);
        const int i0 = meta.i0;
        float* out = a.rev + meta.off;
        for (int j = tid; j < m; j += NT) {
            const int i = i0 + j;
            float2 e = make_float2(-1.f, 0.f);
            v3 bev = c.kinv(mk3(c.rs.x + (float)tZ, c.rs.y + (float)i, c.rs.z + (float)tY));
            if (bev.x >= 0.f && bev.y >= 0.f && bev.z >= 0.f && bev.x < a.bev_size.x && bev.y < a.bev_size.y &&
                bev.z < a.bev_size.z) {
                v3 g = a.binv(bev) + a.off;
                if (g.x >= a.bb0.x && g.x < a.bb1.x && g.y >= a.bb0.y && g.y < a.bb1.y && g.z >= a.bb0.z &&
                    g.z < a.bb1.z) {
                    const float2 dt = tex3_pair(a.dt, a.vd.nx, a.vd.ny, a.vd.nz, g.x + 0.5f, g.y + 0.5f, g.z + 0.5f);
                    e.x = 0.5f * (a.delr * dt.x);
                    e.y = dt.y;
                }
            }
            s_line[j] = e;
        }
        if (tid < nr) {
            sR[tid] = tR;
            sIW[tid] = tIW;
            sF[tid] = tF;
            sB[tid] = tB;
        }
        __syncthreads();
        // valid segment of the line (contiguous: the crop box and the calc_bbox are convex)
        int vlo = m, vhi = -1, nvalid = 0;
        for (int j = tid; j < m; j += NT)
            if (s_line[j].x >= 0.f) {
                vlo = min(vlo, j);
                vhi = max(vhi, j);
                nvalid++;
            }
        for (int o = 32; o > 0; o >>= 1) {
            vlo = min(vlo, __shfl_xor(vlo, o, 64));
            vhi = max(vhi, __shfl_xor(vhi, o, 64));
            nvalid += __shfl_xor(nvalid, o, 64);
        }
        if (NW > 1) {
            if (lane == 0) {
                s_red[wid][0] = vlo;
                s_red[wid][1] = vhi;
                s_red[wid][2] = nvalid;
            }
            __syncthreads();
            vlo = s_red[0][0];
            vhi = s_red[0][1];
            nvalid = s_red[0][2];
            for (int w = 1; w < NW; w++) {
                vlo = min(vlo, s_red[w][0]);
                vhi = max(vhi, s_red[w][1]);
                nvalid += s_red[w][2];
            }
        }
        const float nphi_f = (float)a.nphi;
        for (int j = tid; j < m; j += NT)
            if (j < vlo || j > vhi) out[j] = 0.f;
        const int n = vhi - vlo + 1;
        if (n >= a.ramp_min && nvalid == n) {
            // kernel rows as C(rho) = beta_s + sigma_s * rho on (R_{s-1}, R_s]
            if (tid < 2 * nr) {
                const int d = tid / nr, q = tid - d * nr;
                const float* row = d == 0 ? sF : sB;
                const float prev = q == 0 ? 0.f : row[q - 1];
                const float sg = (row[q] - prev) * sIW[q];
                sSig[d][q] = (double)sg;
                sBeta[d][q] = (double)prev - (double)sg * (q == 0 ? 0.0 : (double)sR[q - 1]);
            }
            const float2* sv = s_line + vlo;
            // radiological centres (FP64 scan) and the two FP64 prefix sums, one chunk per thread
            const int per = (n + NT - 1) / NT;
            const int ca = min(n, tid * per), cb = min(n, ca + per);
            double loc = 0.0;
            for (int k = ca; k < cb; k++) loc += 2.0 * (double)sv[k].x;
            double P = block_excl_sum<NT>(loc, s_tot);
            double le = 0.0, le2 = 0.0;
            for (int k = ca; k < cb; k++) {
                const float2 e = sv[k];
                const float ck = (float)(P + (double)e.x);
                rc[k] = ck;
                P += 2.0 * (double)e.x;
                const double tp = k > 0 ? (double)sv[k - 1].y : 0.0, tn = k + 1 < n ? (double)sv[k + 1].y : 0.0;
                le += ((double)e.y - tp) * (double)ck;
                le2 += ((double)e.y - tn) * (double)ck;
            }
            double e1 = block_excl_sum<NT>(le, s_tot);
            double e2 = block_excl_sum<NT>(le2, s_tot);
            for (int k = ca; k < cb; k++) {
                const double ck = (double)rc[k], t = (double)sv[k].y;
                const double tp = k > 0 ? (double)sv[k - 1].y : 0.0, tn = k + 1 < n ? (double)sv[k + 1].y : 0.0;
                rE[k] = e1;
                rE2[k] = e2;
                e1 += (t - tp) * ck;
                e2 += (t - tn) * ck;
            }
            if (ca < cb && cb == n) {
                rE[n] = e1;
                rE2[n] = e2;
                for (int q = 0; q < 4; q++) rc[n + q] = INFINITY;  // lookups probe up to 3 past n
            }
            __syncthreads();
            // lookup table: lb[u] = first k with c_k >= c0 + u * D, U = 2n cells over [c0, c_{n-1}]
            RampLine L;
            L.c = rc;
            L.E = rE;
            L.E2 = rE2;
            L.s = sv;
            L.lb = rlb;
            L.n = n;
            L.U = 2 * n;
            L.c0 = rc[0];
            L.cn = rc[n - 1];
            const float range = L.cn - L.c0;
            const float D = range > 0.f ? range / (float)L.U : 0.f;
            L.invD = range > 0.f ? (float)L.U / range : 0.f;
            // sample k owns the cells whose threshold lies in (c_{k-1}, c_k] (c_{-1} = -inf, c_n =
            // +inf): each of the U + 2 entries is written by exactly one sample
            for (int k = tid; k <= n; k += NT) {
                const int u0 = k == 0 ? 0 : ramp_cell_after(L.c0, D, L.invD, L.U + 2, rc[k - 1]);
                const int u1 = k == n ? L.U + 2 : ramp_cell_after(L.c0, D, L.invD, L.U + 2, rc[k]);
                for (int u = u0; u < u1; u++) rlb[u] = (uint16_t)k;
            }
            __syncthreads();
            {
                for (int j = tid; j < n; j += NT) {
                    const double v = ramp_back(L, j, sSig[0], sBeta[0], sR, nr, iters) +
                                     ramp_fwd(L, j, sSig[1], sBeta[1], sR, nr, iters);
                    out[vlo + j] = (float)v / nphi_f;
                }
            }
        } else {
            for (int base = vlo; base <= vhi; base += NT) {
                const int j = base + tid;
                const bool act = j <= vhi && s_line[min(j, vhi)].x >= 0.f;
                const float v = walk_both(s_line, min(j, vhi), vlo, vhi, sR, sIW, sF, sB, nr, act, iters);
                if (j <= vhi) out[j] = act ? v / nphi_f : 0.f;
            }
        }
    }
    // every line's slot is written by exactly one of the length-class launches (empty lines' slots
    // by the class-list pass when the launches take lists)
    if (a.counts && (mine || (!a.list && meta.len == 0 && a.len_lo == 1))) {
        for (int o = 32; o > 0; o >>= 1) iters += __shfl_xor(iters, o, 64);
        if (NW == 1) {
            if (tid == 0) a.counts[line] = iters;
        } else {
            if (lane == 0) s_cnt[wid] = iters;
            __syncthreads();
            if (tid == 0) {
                uint32_t t = 0;
                for (int w = 0; w < NW; w++) t += s_cnt[w];
                a.counts[line] = t;
            }
        }
    }
}

constexpr int kPvcsGroup = 4;

struct PvcsArgs {
    const ConeDev* cones;
    const LineMeta* lines;
    const float* rev;
    float* dose;
    int ncones;
    int q0, q1;                 // cones of this pass
    int nx, ny;                 // full volume dims (x, y)
    int x0, y0, z0, sx, sy, sz; // PVCS index box
    int tx, ty, tz;             // tiles per axis
    m33 bfwd;                   // BEV <- PVCS rotation (index units)
    v3 off;                     // BEV origin in PVCS index units
    v3 bev_lo, bev_hi;          // dilated crop box (BEV index units)
};

// One lane per PVCS voxel, 16x4x4 voxel tiles in XCD-aware order.  Voxels outside the dilated
// crop box receive nothing from any cone and leave at once; the others sum every cone's trilinear
// REV dose (compact store, REV2XYZdose's fetch) in a register and add it to the dose volume once.
__global__ void __launch_bounds__(256) pvcs_kernel(PvcsArgs a) {
    __shared__ ConeDev sc[kMaxRays];
    const int tid = threadIdx.x;
    for (int q = tid; q < a.ncones; q += 256) sc[q] = a.cones[q];
    __syncthreads();
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const int txi = (int)(tile % (uint32_t)a.tx);
    const int tyi = (int)((tile / (uint32_t)a.tx) % (uint32_t)a.ty);
    const int tzi = (int)(tile / ((uint32_t)a.tx * (uint32_t)a.ty));
    const int X = txi * 16 + (tid & 15), Y = tyi * 4 + ((tid >> 4) & 3), Z = tzi * 4 + (tid >> 6);
    if (X >= a.sx || Y >= a.sy || Z >= a.sz) return;
    const int gx = a.x0 + X, gy = a.y0 + Y, gz = a.z0 + Z;
    const v3 g = mk3((float)gx, (float)gy, (float)gz);
    const v3 bv = a.bfwd(g - a.off);
    if (!(bv.x >= a.bev_lo.x && bv.y >= a.bev_lo.y && bv.z >= a.bev_lo.z && bv.x < a.bev_hi.x && bv.y < a.bev_hi.y &&
          bv.z < a.bev_hi.z))
        return;
    // Cones in groups of kPvcsGroup: the line-table loads of a group are issued together, then its
    // REV loads (branch-free: taps outside a line's stored interval read sample 0 of the store and
    // are replaced by zero), so each lane keeps several dependent gathers in flight.  The
    // arithmetic per cone is rev_fetch's and the sum runs in cone order.  The voxel's dose is read
    // first so that its latency overlaps the gathers (this lane alone writes the voxel).
    float* dst = a.dose + ((size_t)gx + (size_t)a.nx * ((size_t)gy + (size_t)a.ny * gz));
    const float old = *dst;
    float sum = 0.f;
    for (int q0 = a.q0; q0 < a.q1; q0 += kPvcsGroup) {
        LineMeta mt[kPvcsGroup][4];
        float fa[kPvcsGroup], fb[kPvcsGroup], fc[kPvcsGroup];
        int di[kPvcsGroup];
        bool ok[kPvcsGroup];
#pragma unroll
        for (int u = 0; u < kPvcsGroup; u++) {
            ok[u] = false;
            di[u] = 0;
            fa[u] = fb[u] = fc[u] = 0.f;
#pragma unroll
            for (int t = 0; t < 4; t++) mt[u][t] = LineMeta{0u, 0, 0};
            const int q = q0 + u;
            if (q >= a.q1) continue;
            const ConeDev& c = sc[q];
            const v3 r = c.a(g) + c.t;
            if (r.y >= (float)c.nx || r.z >= (float)c.ny || r.x >= (float)c.nz || r.x < 0.f || r.y < 0.f || r.z < 0.f)
                continue;
            const float ub = (r.y + 0.5f) - 0.5f, vb = (r.z + 0.5f) - 0.5f, wb = (r.x + 0.5f) - 0.5f;
            if (!(ub > -1.f && vb > -1.f && wb > -1.f && ub < (float)c.nx && vb < (float)c.ny && wb < (float)c.nz))
                continue;
            const float fu = floorf(ub), fv = floorf(vb), fw = floorf(wb);
            const int i = (int)fu, j = (int)fv, k = (int)fw;
            ok[u] = true;
            di[u] = i;
            fa[u] = ub - fu;
            fb[u] = vb - fv;
            fc[u] = wb - fw;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int jj = j + (t & 1), kk = k + (t >> 1);
                if (jj >= 0 && jj < c.ny && kk >= 0 && kk < c.nz)
                    mt[u][t] = a.lines[c.line0 + (uint64_t)jj + (uint64_t)c.ny * kk];
            }
        }
#pragma unroll
        for (int u = 0; u < kPvcsGroup; u++) {
            if (!ok[u]) continue;
            float tv[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const LineMeta m = mt[u][t];
                const int d = di[u] - m.i0;
                const bool vlo = d >= 0 && d < m.len, vhi = d + 1 >= 0 && d + 1 < m.len;
                const float x0 = a.rev[vlo ? m.off + (uint32_t)d : 0u];
                const float x1 = a.rev[vhi ? m.off + (uint32_t)(d + 1) : 0u];
                const float lo = vlo ? x0 : 0.f, hi = vhi ? x1 : 0.f;
                tv[t] = lo + fa[u] * (hi - lo);
            }
            const float e0 = tv[0] + fb[u] * (tv[1] - tv[0]);
            const float e1 = tv[2] + fb[u] * (tv[3] - tv[2]);
            sum += e0 + fc[u] * (e1 - e0);
        }
    }
    if (sum != 0.f) *dst = old + sum;
}

// Interleaves density and TERMA over the CT grid for the cone kernel's paired fetch; TERMA is 0
// outside the calc_bbox, which is what the bbox-bounded TERMA fetch returns there.
__global__ void __launch_bounds__(256) pack_dt_kernel(const float* __restrict__ dens, const float* __restrict__ terma,
                                                      int nx, int ny, int nz, int bx0, int by0, int bz0, int bx, int by,
                                                      int bz, float2* __restrict__ out) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= nx || y >= ny) return;
    const size_t v = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
    const int tx = x - bx0, ty = y - by0, tz = z - bz0;
    const float t = (tx >= 0 && ty >= 0 && tz >= 0 && tx < bx && ty < by && tz < bz)
                        ? terma[(size_t)tx + (size_t)bx * ((size_t)ty + (size_t)by * tz)]
                        : 0.f;
    out[v] = make_float2(dens[v], t);
}

// ---- structure projection onto fluence maps (preprocess: findFluenceProjection + cudaRaycast,
// dosecalc-preprocess/cudaMemFunctions.cu:308-393, kernels.cuh:136-224) ---------------------------
struct ProjBeam {
    v3 src, iso;
    Rot2 inv;         // FCS -> PVCS (inverseRotateBeamAtOriginRHS)
    int fx, fy;       // fluence map size
    float bsx, bsy;   // beamlet size [cm]
    uint32_t off;     // first beamlet of this beam in the output
};

// One lane per beamlet of any beam.  Nine sub-rays (offsets -1/0/+1 half beamlets along each
// fluence axis) run from the source to three times the source-to-cell distance through the
// trilinearly sampled mask with the TERMA path's Siddon march; the beamlet is 1 when any sub-ray
// accumulates a positive path integral.
__global__ void __launch_bounds__(64) projection_kernel(const float* __restrict__ mask, VolDesc vd,
                                                        const ProjBeam* __restrict__ beams, int nbeams, uint32_t total,
                                                        float* __restrict__ out) {
    const uint32_t t = blockIdx.x * 64 + threadIdx.x;
    if (t >= total) return;
    int lo = 0, hi = nbeams - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (beams[mid].off <= t) lo = mid; else hi = mid - 1;
    }
    const ProjBeam& b = beams[lo];
    const int pix = (int)(t - b.off);
    const int rx = pix % b.fx, rz = pix / b.fx;
    bool hit = false;
    for (int sr = 0; sr < 9 && !hit; sr++) {
        const int sx = sr % 3 - 1, sz = sr / 3 - 1;
        const v3 fcs = mk3(b.bsx * ((float)rx + 0.5f * (float)sx - (float)b.fx * 0.5f + 0.5f), 0.f,
                           b.bsy * ((float)rz + 0.5f * (float)sz - (float)b.fy * 0.5f + 0.5f));
        v3 ctr = rot2(b.inv, fcs) + b.iso;
        ctr = ctr + 1e-9f;
        const v3 sd = ctr - b.src;
        const v3 sink = b.src + 3.f * sd;
        const float rpl = siddon_march(sink, vd.start, vd.end, vd.voxel, b.src, mask, vd.nx, vd.ny, vd.nz,
                                       [](float, float, float, float) {}, nullptr);
        hit = rpl > 0.f;
    }
    out[t] = hit ? 1.f : 0.f;
}

// ---- fluence-map convolution (the FCBB lateral step; no reference source) --------------------
// out[y][x] = sum_j sum_i K[j][i] * F[y - (j - kh/2)][x - (i - kw/2)], zero outside the map, the
// terms added j outer, i inner.  One lane per output of a 16x16 tile; the tile plus its halo and
// the kernel sit in LDS.
constexpr int kConvTile = 16, kConvMaxK = 31;

__global__ void __launch_bounds__(256) fluence_conv_kernel(const float* __restrict__ f, int w, int h,
                                                           const float* __restrict__ k, int kw, int kh,
                                                           float* __restrict__ out) {
    __shared__ float s_k[kConvMaxK * kConvMaxK];
    __shared__ float s_f[(kConvTile + kConvMaxK - 1) * (kConvTile + kConvMaxK - 1)];
    const int map = blockIdx.z;
    const float* F = f + (size_t)map * w * h;
    const int x0 = blockIdx.x * kConvTile, y0 = blockIdx.y * kConvTile;
    const int rx = kw / 2, ry = kh / 2;
    const int sw = kConvTile + kw - 1, sh = kConvTile + kh - 1;
    const int tid = threadIdx.x;
    for (int q = tid; q < kw * kh; q += 256) s_k[q] = k[q];
    // s_f[v][u] = F[y0 + ry - v][x0 + rx - u]: the flipped window, so output (tx, ty) reads
    // s_f[ty + j][tx + i] for kernel tap (i, j)
    for (int q = tid; q < sw * sh; q += 256) {
        const int u = q % sw, v = q / sw;
        const int xx = x0 + (kConvTile - 1) - u + rx, yy = y0 + (kConvTile - 1) - v + ry;
        s_f[q] = (xx >= 0 && yy >= 0 && xx < w && yy < h) ? F[(size_t)yy * w + xx] : 0.f;
    }
    __syncthreads();
    const int tx = tid % kConvTile, ty = tid / kConvTile;
    const int x = x0 + tx, y = y0 + ty;
    if (x >= w || y >= h) return;
    // output (x, y) needs F[y - j + ry][x - i + rx] = s_f[(kConvTile-1-ty) + j][(kConvTile-1-tx) + i]
    const int bu = kConvTile - 1 - tx, bv = kConvTile - 1 - ty;
    float acc = 0.f;
    for (int j = 0; j < kh; j++)
        for (int i = 0; i < kw; i++) acc += s_k[j * kw + i] * s_f[(bv + j) * sw + bu + i];
    out[(size_t)map * w * h + (size_t)y * w + x] = acc;
}

__global__ void mask_to_float_kernel(const uint8_t* __restrict__ m, float* __restrict__ f, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        f[i] = (float)m[i];
}

__global__ void add_kernel(float* __restrict__ dst, const float* __restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] += src[i];
}

// ------------------------------------------------------------------------------------------------
// host pipeline

namespace {

inline size_t bbox_voxels(const cccs_ctx* ctx) {
    return (size_t)ctx->prob.bbox_size[0] * ctx->prob.bbox_size[1] * ctx->prob.bbox_size[2];
}
inline size_t vol_voxels(const cccs_ctx* ctx) {
    return (size_t)ctx->prob.size[0] * ctx->prob.size[1] * ctx->prob.size[2];
}

TermaArgs terma_args(cccs_ctx* ctx, const cccs_beam& b, const float* d_fmap, float* d_terma, int beamletnum) {
    TermaArgs a;
    a.terma = d_terma;
    a.fmap = d_fmap;
    a.dens = ctx->d_dens;
    a.vd = ctx->vol;
    a.beam = make_beam_dev(b);
    a.spec = ctx->spec;
    a.bx = (int)ctx->prob.bbox_size[0]; a.by = (int)ctx->prob.bbox_size[1]; a.bz = (int)ctx->prob.bbox_size[2];
    a.ox = (int)ctx->prob.bbox_start[0]; a.oy = (int)ctx->prob.bbox_start[1]; a.oz = (int)ctx->prob.bbox_start[2];
    a.ss = std::max(1u, ctx->prob.ss_factor);
    a.bhc = ctx->prob.beamhard_correct;
    a.beamletnum = beamletnum;
    a.counts = nullptr;
    return a;
}

int launch_terma(cccs_ctx* ctx, const TermaArgs& a0) {
    dim3 blk(64, 4, 1), grd((a0.bx + 63) / 64, (a0.by + 3) / 4, a0.bz);
    const size_t blocks = (size_t)grd.x * grd.y * grd.z;
    TermaArgs a = a0;
    a.counts = prof_counter(ctx, blocks);
    const double nbb = (double)a.bx * a.by * a.bz;
    const double nct = (double)a.vd.nx * a.vd.ny * a.vd.nz;
    int tok = prof_begin(ctx, K_TERMA);
    hipLaunchKernelGGL(terma_kernel, grd, blk, 0, ctx->stream, a);
    prof_end(ctx, tok, 4.0 * nct + 4.0 * nbb, 0.0, (uint64_t)nbb);
    CCCS_HIP(hipGetLastError());
    return CCCS_OK;
}

const float* angle_cache_lookup(const cccs_ctx* ctx, const cccs_beam& b);

// TERMA from the resident unit-TERMA grid of the beam's source (angle cache).
int launch_terma_from_unit(cccs_ctx* ctx, const TermaArgs& a, const float* unit) {
    dim3 blk(64, 4, 1), grd((a.bx + 63) / 64, (a.by + 3) / 4, a.bz);
    const double nbb = (double)a.bx * a.by * a.bz;
    int tok = prof_begin(ctx, K_TERMA_CACHED);
    hipLaunchKernelGGL(terma_from_unit_kernel, grd, blk, 0, ctx->stream, a, unit);
    prof_end(ctx, tok, 8.0 * nbb, 0.0, (uint64_t)nbb);
    CCCS_HIP(hipGetLastError());
    return CCCS_OK;
}

// Crop box + REV grids (nvbbRayConvolve_host.cu:429-587), from the reduced extent.
int beam_geometry(cccs_ctx* ctx, const cccs_beam& b, const float* red, BeamGeom& g) {
    const cccs_problem& p = ctx->prob;
    // the reference starts max at FLT_MIN and min at FLT_MAX (:436-437)
    v3 bmin = mk3(std::min(red[0], FLT_MAX), std::min(red[1], FLT_MAX), std::min(red[2], FLT_MAX));
    v3 bmax = mk3(std::max(red[3], FLT_MIN), std::max(red[4], FLT_MIN), std::max(red[5], FLT_MIN));
    if (red[6] < 0.5f) return set_error(CCCS_EGEOM, "beam has no non-zero TERMA inside the calc_bbox");
    bmin.x -= p.penumbra / p.rev_latspacing;
    bmin.y -= p.penumbra / p.rev_longspacing;
    bmin.z -= p.penumbra / p.rev_latspacing;
    bmax.x += p.penumbra / p.rev_latspacing;
    bmax.y += p.penumbra / p.rev_longspacing;
    bmax.z += p.penumbra / p.rev_latspacing;
    float sx = std::ceil(std::fabs(bmax.x - bmin.x)), sy = std::ceil(std::fabs(bmax.y - bmin.y)),
          sz = std::ceil(std::fabs(bmax.z - bmin.z));
    if (!(sx >= 1.f && sx < 1000.f && sy >= 1.f && sy < 1000.f && sz >= 1.f && sz < 1000.f))
        return set_error(CCCS_EGEOM, "BEV crop box size out of range (0, 1000)");
    g.bev_size[0] = (uint32_t)sx; g.bev_size[1] = (uint32_t)sy; g.bev_size[2] = (uint32_t)sz;
    g.bev_min = bmin;
    g.bev_max = bmax;
    dm33 binv = dm_beam_inv(b.azimuth, b.zenith, b.coll);
    v3 tmin = rot2(rot2_beam_inv(b.azimuth, b.zenith, b.coll), bmin);
    g.lim_min = mk3(p.start[0] + (tmin.x + (float)p.bbox_start[0]) * p.voxel[0],
                    p.start[1] + (tmin.y + (float)p.bbox_start[1]) * p.voxel[1],
                    p.start[2] + (tmin.z + (float)p.bbox_start[2]) * p.voxel[2]);
    // REV grid per cone: bounding box of the rotated crop box (cm), sized in spacing units
    v3 lim = mk3(sx * p.rev_latspacing, sy * p.rev_longspacing, sz * p.rev_latspacing);
    g.nrays = ctx->nrays;
    g.total_lines = 0;
    g.total_rev = 0;
    for (int r = 0; r < ctx->nrays; r++) {
        float th = ctx->theta[r] * 3.14159265358979323846f / 180.f;
        float ph = ctx->phi[r] * 3.14159265358979323846f / 180.f;
        Rot2 kf = rot2_kernel_fwd(th, ph);
        v3 mn = mk3(FLT_MAX, FLT_MAX, FLT_MAX), mx = mk3(FLT_MIN, FLT_MIN, FLT_MIN);
        for (int c = 0; c < 8; c++) {
            v3 in = mk3(lim.x * (float)(c & 1), lim.y * (float)((c >> 1) & 1), lim.z * (float)((c >> 2) & 1));
            v3 o = rot2(kf, in);
            if (o.x < mn.x) mn.x = o.x;
            if (o.y < mn.y) mn.y = o.y;
            if (o.z < mn.z) mn.z = o.z;
            if (o.x > mx.x) mx.x = o.x;
            if (o.y > mx.y) mx.y = o.y;
            if (o.z > mx.z) mx.z = o.z;
        }
        uint32_t n0 = (uint32_t)std::ceil(std::fabs(mx.y - mn.y) / p.rev_longspacing);
        uint32_t n1 = (uint32_t)std::ceil(std::fabs(mx.z - mn.z) / p.rev_latspacing);
        uint32_t n2 = (uint32_t)std::ceil(std::fabs(mx.x - mn.x) / p.rev_latspacing);
        if (n0 > ctx->max_rev_beam[0] || n1 > ctx->max_rev_beam[1] || n2 > ctx->max_rev_beam[2])
            return set_error(CCCS_EGEOM, "REV grid exceeds max_rev_size (ArraySizeError)");
        if (n0 > (uint32_t)kMaxLine) return set_error(CCCS_EGEOM, "REV line longer than 1024 samples");
        g.rev_min[r] = mn;
        g.rev_size[r][0] = n0; g.rev_size[r][1] = n1; g.rev_size[r][2] = n2;
        g.total_lines += (uint64_t)n1 * n2;
        g.total_rev += (uint64_t)n0 * n1 * n2;
    }
    // PVCS index box reachable from the crop box dilated by 2 index units (trilinear support of
    // a REV sample is within sqrt(3) of it), clamped to the calc_bbox
    v3 off = mk3((g.lim_min.x - p.start[0]) / p.voxel[0], (g.lim_min.y - p.start[1]) / p.voxel[1],
                 (g.lim_min.z - p.start[2]) / p.voxel[2]);
    double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
    for (int c = 0; c < 8; c++) {
        v3 cb = mk3((c & 1) ? sx + 2.f : -2.f, ((c >> 1) & 1) ? sy + 2.f : -2.f, ((c >> 2) & 1) ? sz + 2.f : -2.f);
        v3 q = dm_apply_f(binv, cb) + off;
        lo[0] = std::min(lo[0], (double)q.x); lo[1] = std::min(lo[1], (double)q.y); lo[2] = std::min(lo[2], (double)q.z);
        hi[0] = std::max(hi[0], (double)q.x); hi[1] = std::max(hi[1], (double)q.y); hi[2] = std::max(hi[2], (double)q.z);
    }
    for (int k = 0; k < 3; k++) {
        long l0 = (long)std::floor(lo[k]) - 1, h0 = (long)std::ceil(hi[k]) + 2;
        long b0 = (long)p.bbox_start[k], b1 = b0 + (long)p.bbox_size[k];
        g.pvcs_lo[k] = (int)std::max(b0, std::min(b1, l0));
        g.pvcs_hi[k] = (int)std::max(b0, std::min(b1, h0));
    }
    return CCCS_OK;
}

// Device and host scratch of one full-beam dose computation: the context's grow-only buffers, or,
// for the no-allocation entry point, consecutive 256-byte aligned pieces of the library's fixed
// device array with host staging on the caller's stack.
struct BeamScratch {
    cccs_ctx* ctx = nullptr;
    char* arena = nullptr;          // fixed device array, or nullptr: the context's buffers
    size_t cap = 0, used = 0;
    float* host_small = nullptr;    // >= 64 bytes of host staging
    ConeDev* host_cones = nullptr;  // kMaxRays entries of host staging
    int get(DevBuf& buf, size_t bytes, void** out) {
        if (!arena) {
            CCCS_TRY(buf.ensure(bytes));
            *out = buf.p;
            return CCCS_OK;
        }
        const size_t at = (used + 255) & ~size_t(255);
        if (bytes > cap || at > cap - bytes)
            return set_error(CCCS_ENOMEM, "beam needs more than the " + std::to_string(cap >> 20) +
                                              " MiB fixed scratch of the no-allocation entry point");
        used = at + bytes;
        *out = arena + at;
        return CCCS_OK;
    }
    template <class T>
    int get(DevBuf& buf, size_t count, T** out) {
        void* v = nullptr;
        CCCS_TRY(get(buf, sizeof(T) * std::max<size_t>(1, count), &v));
        *out = static_cast<T*>(v);
        return CCCS_OK;
    }
};

// Scratch from the context's own (allocating) buffers.
int context_scratch(cccs_ctx* ctx, BeamScratch& S) {
    S = BeamScratch();
    S.ctx = ctx;
    CCCS_TRY(ctx->h_small.ensure(64));
    CCCS_TRY(ctx->h_cones.ensure(sizeof(ConeDev) * kMaxRays));
    S.host_small = ctx->h_small.as<float>();
    S.host_cones = ctx->h_cones.as<ConeDev>();
    return CCCS_OK;
}

int run_crop(BeamScratch& S, const float* d_terma, const cccs_beam& b, float* red_host) {
    cccs_ctx* ctx = S.ctx;
    float* part = nullptr;
    CCCS_TRY(S.get(ctx->partial, kCropBlocks * 8 + 8, &part));
    Rot2 rot = rot2_beam_fwd(b.azimuth, b.zenith, b.coll);
    const cccs_problem& p = ctx->prob;
    size_t n = bbox_voxels(ctx);
    int blocks = (int)std::min<size_t>(kCropBlocks, (n + 255) / 256);
    int tok = prof_begin(ctx, K_CROP);
    hipLaunchKernelGGL(crop_partial_kernel, dim3(blocks), dim3(256), 0, ctx->stream, d_terma, (int)p.bbox_size[0],
                       (int)p.bbox_size[1], (int)p.bbox_size[2], rot, part);
    hipLaunchKernelGGL(crop_final_kernel, dim3(1), dim3(256), 0, ctx->stream, part, blocks, part + kCropBlocks * 8);
    prof_end(ctx, tok, 4.0 * (double)n, 0.0, n);
    CCCS_HIP(hipGetLastError());
    CCCS_HIP(hipMemcpyAsync(S.host_small, part + kCropBlocks * 8, 7 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(red_host, S.host_small, 7 * sizeof(float));
    return CCCS_OK;
}

// Everything after TERMA: geometry, cones, PVCS accumulation into d_dose.
// ---- per-class line lists -----------------------------------------------------------------------
// The cone kernel runs one launch per stored-length class (LDS sized for the class's longest line).
// Each launch takes the list of its own lines, in line order, so that no workgroup is spent on an
// empty line or on a line of another class.  Three passes: per-block class counts, one block that
// turns them into list offsets (class by class), and a stable scatter of the line indices.
constexpr int kClassBlock = 1024;

struct ClassArgs {
    const LineMeta* lines;
    uint64_t n;
    int hi[kConeClasses];  // longest stored line of each class (ascending)
    int ncls;
    uint32_t* bcnt;        // [blocks][kConeClasses]: counts, then the blocks' list offsets
    uint32_t* tot;         // [kConeClasses]: lines per class
    ConeLine* list;        // lines with their cone and record, class by class
    const ConeDev* cones;  // the beam's cones (the entries' cone indices)
    int ncones;
    uint32_t* counts;      // profiling: the slots of empty lines are zeroed here (or nullptr)
};

__device__ __forceinline__ int class_of_line(const ClassArgs& a, int len) {
    if (len <= 0) return -1;
    for (int c = 0; c < a.ncls - 1; c++)
        if (len <= a.hi[c]) return c;
    return a.ncls - 1;
}

__global__ void __launch_bounds__(kClassBlock) class_count_kernel(ClassArgs a) {
    __shared__ uint32_t s[kClassBlock / 64][kConeClasses];
    const uint64_t i = (uint64_t)blockIdx.x * kClassBlock + threadIdx.x;
    const int cls = i < a.n ? class_of_line(a, a.lines[i].len) : -1;
    if (a.counts && i < a.n && cls < 0) a.counts[i] = 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c = 0; c < kConeClasses; c++) {
        const uint64_t m = __ballot(cls == c);
        if (lane == 0) s[w][c] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < kConeClasses) {
        uint32_t t = 0;
        for (int q = 0; q < kClassBlock / 64; q++) t += s[q][threadIdx.x];
        a.bcnt[(size_t)blockIdx.x * kConeClasses + threadIdx.x] = t;
    }
}

// One workgroup: exclusive scan of the block counts of each class, the classes one after another.
__global__ void __launch_bounds__(kClassBlock) class_scan_kernel(ClassArgs a, uint32_t nb) {
    __shared__ uint32_t s[kClassBlock];
    const int t = threadIdx.x;
    uint32_t base = 0;
    for (int c = 0; c < kConeClasses; c++) {
        const uint32_t start = base;
        for (uint32_t b0 = 0; b0 < nb; b0 += kClassBlock) {
            const uint32_t b = b0 + t;
            const uint32_t v = b < nb ? a.bcnt[(size_t)b * kConeClasses + c] : 0u;
            s[t] = v;
            __syncthreads();
            for (int off = 1; off < kClassBlock; off <<= 1) {
                const uint32_t add = t >= off ? s[t - off] : 0u;
                __syncthreads();
                s[t] += add;
                __syncthreads();
            }
            if (b < nb) a.bcnt[(size_t)b * kConeClasses + c] = base + s[t] - v;
            base += s[kClassBlock - 1];
            __syncthreads();
        }
        if (t == 0) a.tot[c] = base - start;
    }
}

__global__ void __launch_bounds__(kClassBlock) class_scatter_kernel(ClassArgs a) {
    __shared__ uint32_t s[kClassBlock / 64][kConeClasses];
    const uint64_t i = (uint64_t)blockIdx.x * kClassBlock + threadIdx.x;
    const int cls = i < a.n ? class_of_line(a, a.lines[i].len) : -1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    uint32_t rank = 0;
    for (int c = 0; c < kConeClasses; c++) {
        const uint64_t m = __ballot(cls == c);
        if (lane == 0) s[w][c] = (uint32_t)__popcll(m);
        if (cls == c) rank = (uint32_t)__popcll(m & below);
    }
    __syncthreads();
    if (cls >= 0) {
        uint32_t off = a.bcnt[(size_t)blockIdx.x * kConeClasses + cls] + rank;
        for (int q = 0; q < w; q++) off += s[q][cls];
        ConeLine e;
        e.line = (uint32_t)i;
        e.cone = (uint32_t)cone_of_line(a.cones, a.ncones, i);
        e.meta = a.lines[i];
        a.list[off] = e;
    }
}

void launch_cone(int threads, dim3 grid, size_t lds, hipStream_t st, const ConeArgs& ca) {
    switch (threads) {
        case 256: hipLaunchKernelGGL(cone_kernel<256>, grid, dim3(256), lds, st, ca); break;
        case 128: hipLaunchKernelGGL(cone_kernel<128>, grid, dim3(128), lds, st, ca); break;
        default: hipLaunchKernelGGL(cone_kernel<64>, grid, dim3(64), lds, st, ca); break;
    }
}

int run_cones(BeamScratch& S, const cccs_beam& b, const float* d_terma, float* d_dose, BeamGeom* gout) {
    cccs_ctx* ctx = S.ctx;
    float red[7];
    CCCS_TRY(run_crop(S, d_terma, b, red));
    BeamGeom g;
    CCCS_TRY(beam_geometry(ctx, b, red, g));
    if (gout) *gout = g;
    if (!d_dose) return CCCS_OK;
    const cccs_problem& p = ctx->prob;
    if (g.total_lines > 0x7fffffffULL || g.total_rev > 0xffffffffULL)
        return set_error(CCCS_EGEOM, "REV grids exceed 2^32 samples");
    ConeDev* d_cones = nullptr;
    LineMeta* d_lines = nullptr;
    CCCS_TRY(S.get(ctx->cones, kMaxRays, &d_cones));
    CCCS_TRY(S.get(ctx->lines, g.total_lines, &d_lines));
    ConeDev* hc = S.host_cones;
    dm33 bfwd = dm_beam_fwd(b.azimuth, b.zenith, b.coll);
    dm33 binv = dm_beam_inv(b.azimuth, b.zenith, b.coll);
    v3 rv = mk3(p.rev_latspacing, p.rev_longspacing, p.rev_latspacing);
    v3 off = mk3((g.lim_min.x - p.start[0]) / p.voxel[0], (g.lim_min.y - p.start[1]) / p.voxel[1],
                 (g.lim_min.z - p.start[2]) / p.voxel[2]);
    uint64_t line0 = 0, roff = 0;
    const int half = (int)p.ntheta / 2;
    for (int r = 0; r < g.nrays; r++) {
        float th = ctx->theta[r] * 3.14159265358979323846f / 180.f;
        float ph = ctx->phi[r] * 3.14159265358979323846f / 180.f;
        dm33 kf = dm_kernel_fwd(th, ph), ki = dm_kernel_inv(th, ph);
        ConeDev& c = hc[r];
        c.kinv = to_f(ki);
        dm33 a = dm_mul(kf, bfwd);
        c.a = to_f(a);
        c.rs = mk3(g.rev_min[r].x / rv.x, g.rev_min[r].y / rv.y, g.rev_min[r].z / rv.z);
        // REV coordinate of PVCS index 0: a * (-off) - rs
        v3 t0 = dm_apply_f(a, mk3(-off.x, -off.y, -off.z));
        c.t = mk3(t0.x - c.rs.x, t0.y - c.rs.y, t0.z - c.rs.z);
        c.nx = (int)g.rev_size[r][0];
        c.ny = (int)g.rev_size[r][1];
        c.nz = (int)g.rev_size[r][2];
        c.rowf = r % half;
        c.rowb = (int)p.ntheta - 1 - (r % half);
        c.line0 = line0;
        c.off = roff;
        line0 += (uint64_t)c.ny * c.nz;
        roff += (uint64_t)c.nx * c.ny * c.nz;
    }
    CCCS_HIP(hipMemcpyAsync(d_cones, hc, sizeof(ConeDev) * g.nrays, hipMemcpyHostToDevice, ctx->stream));
    // the host staging may be the caller's stack (no-allocation entry point): the copy completes first
    if (S.arena) CCCS_HIP(hipStreamSynchronize(ctx->stream));
    const size_t nvox = vol_voxels(ctx);
    if (nvox >= (size_t(1) << 31)) return set_error(CCCS_EGEOM, "CT grid beyond 2^31 voxels");
    float2* d_dt = nullptr;
    CCCS_TRY(S.get(ctx->dt, nvox, &d_dt));
    ConeArgs ca;
    ca.list = nullptr;
    ca.dt = d_dt;
    ca.tables = ctx->d_tables;
    ca.cones = d_cones;
    ca.lines = d_lines;
    ca.rev = nullptr;
    ca.nlines = g.total_lines;
    ca.ncones = g.nrays;
    ca.ntheta = (int)p.ntheta;
    ca.nradii = (int)p.nradii;
    ca.nphi = (int)p.nphi;
    ca.delr = p.rev_longspacing;
    ca.counts = nullptr;
    ca.vd = ctx->vol;
    ca.binv = to_f(binv);
    ca.off = off;
    ca.bev_size = mk3((float)g.bev_size[0], (float)g.bev_size[1], (float)g.bev_size[2]);
    ca.bb0 = ldu3(p.bbox_start);
    ca.bb1 = ca.bb0 + ldu3(p.bbox_size);
    ca.bbx = (int)p.bbox_size[0]; ca.bby = (int)p.bbox_size[1]; ca.bbz = (int)p.bbox_size[2];
    if (g.total_lines == 0) return CCCS_OK;
    {
        dim3 blk(64, 4, 1), grd((unsigned)((p.size[0] + 63) / 64), (unsigned)((p.size[1] + 3) / 4), p.size[2]);
        hipLaunchKernelGGL(pack_dt_kernel, grd, blk, 0, ctx->stream, (const float*)ctx->d_dens, d_terma,
                           (int)p.size[0], (int)p.size[1], (int)p.size[2], (int)p.bbox_start[0],
                           (int)p.bbox_start[1], (int)p.bbox_start[2], ca.bbx, ca.bby, ca.bbz, d_dt);
        CCCS_HIP(hipGetLastError());
    }
    // pass 1: stored interval of every line, then offsets into the compact store
    hipLaunchKernelGGL(line_interval_kernel, dim3((unsigned)((g.total_lines + 255) / 256)), dim3(256), 0, ctx->stream,
                       ca);
    CCCS_HIP(hipGetLastError());
    uint32_t* d_scan = nullptr;
    CCCS_TRY(S.get(ctx->scan, line_scan_words(g.total_lines), &d_scan));
    CCCS_TRY(launch_line_scan(ctx, ca.lines, g.total_lines, d_scan));
    // stored-length classes: one launch each, with LDS for the class's longest line, so that the
    // many short lines are not held to the occupancy of the longest
    uint32_t maxn = 1;
    for (int r = 0; r < g.nrays; r++) maxn = std::max(maxn, g.rev_size[r][0]);
    int cls_lo[kConeClasses], cls_hi[kConeClasses], cls_threads[kConeClasses], ncls = 0;
    for (int ci = 0, lo = 1; ci < kConeClasses && lo <= (int)maxn; ci++) {
        const int hi = ci == kConeClasses - 1 ? (int)maxn : std::min<int>(ctx->sched.cone_class_len[ci], (int)maxn);
        if (hi < lo) continue;
        cls_lo[ncls] = lo;
        cls_hi[ncls] = hi;
        cls_threads[ncls] = ctx->sched.cone_class_threads[ci];
        ncls++;
        lo = hi + 1;
    }
    ca.counts = prof_counter(ctx, g.total_lines);
    ClassArgs cl;
    cl.lines = ca.lines;
    cl.n = g.total_lines;
    for (int c = 0; c < kConeClasses; c++) cl.hi[c] = c < ncls ? cls_hi[c] : 0;
    cl.ncls = ncls;
    cl.counts = ca.counts;
    cl.cones = d_cones;
    cl.ncones = g.nrays;
    const uint64_t ncb = (g.total_lines + kClassBlock - 1) / kClassBlock;
    CCCS_TRY(S.get(ctx->cls_cnt, (size_t)ncb * kConeClasses + kConeClasses, &cl.bcnt));
    cl.tot = cl.bcnt + (size_t)ncb * kConeClasses;
    CCCS_TRY(S.get(ctx->cls_list, g.total_lines, &cl.list));
    hipLaunchKernelGGL(class_count_kernel, dim3((unsigned)ncb), dim3(kClassBlock), 0, ctx->stream, cl);
    hipLaunchKernelGGL(class_scan_kernel, dim3(1), dim3(kClassBlock), 0, ctx->stream, cl, (uint32_t)ncb);
    hipLaunchKernelGGL(class_scatter_kernel, dim3((unsigned)ncb), dim3(kClassBlock), 0, ctx->stream, cl);
    CCCS_HIP(hipGetLastError());
    CCCS_HIP(hipMemcpyAsync(S.host_small, ca.lines + (g.total_lines - 1), sizeof(LineMeta), hipMemcpyDeviceToHost,
                            ctx->stream));
    CCCS_HIP(hipMemcpyAsync(S.host_small + 4, cl.tot, sizeof(uint32_t) * kConeClasses, hipMemcpyDeviceToHost,
                            ctx->stream));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    LineMeta last;
    std::memcpy(&last, S.host_small, sizeof(LineMeta));
    uint32_t cls_n[kConeClasses];
    std::memcpy(cls_n, S.host_small + 4, sizeof(cls_n));
    const uint64_t stored = (uint64_t)last.off + (uint64_t)last.len;
    CCCS_TRY(S.get(ctx->rev, stored, &ca.rev));
    // pass 2: per-line resample + walks into the compact store, one workgroup per listed line
    ca.ramp_min = ctx->sched.cone_ramp_min;
    int tok = prof_begin(ctx, K_CONE);
    uint64_t list_off = 0;
    for (int c = 0; c < ncls; c++) {
        if (cls_n[c] == 0) continue;
        ca.len_lo = cls_lo[c];
        ca.len_hi = cls_hi[c];
        ca.maxn = ca.len_hi;
        ca.list = cl.list + list_off;
        list_off += cls_n[c];
        const size_t lds = cone_lds_bytes((uint32_t)ca.maxn, ca.ramp_min <= ca.maxn);
        launch_cone(cls_threads[c], dim3(cls_n[c]), lds, ctx->stream, ca);
        CCCS_HIP(hipGetLastError());
    }
    // algorithmic traffic: density + TERMA in, cone dose out, per stored sample
    prof_end(ctx, tok, 12.0 * (double)stored, 0.0, stored);
    PvcsArgs pa;
    pa.cones = d_cones;
    pa.lines = ca.lines;
    pa.rev = ca.rev;
    pa.dose = d_dose;
    pa.ncones = g.nrays;
    pa.nx = (int)p.size[0];
    pa.ny = (int)p.size[1];
    pa.x0 = g.pvcs_lo[0]; pa.y0 = g.pvcs_lo[1]; pa.z0 = g.pvcs_lo[2];
    pa.sx = g.pvcs_hi[0] - g.pvcs_lo[0]; pa.sy = g.pvcs_hi[1] - g.pvcs_lo[1]; pa.sz = g.pvcs_hi[2] - g.pvcs_lo[2];
    if (pa.sx > 0 && pa.sy > 0 && pa.sz > 0) {
        pa.tx = (pa.sx + 15) / 16;
        pa.ty = (pa.sy + 3) / 4;
        pa.tz = (pa.sz + 3) / 4;
        pa.bfwd = to_f(bfwd);
        pa.off = off;
        // a voxel gets dose only within the trilinear support (< sqrt(3) index units) of a stored
        // sample, and stored samples lie inside the crop box
        pa.bev_lo = mk3(-2.f, -2.f, -2.f);
        pa.bev_hi = mk3((float)g.bev_size[0] + 2.f, (float)g.bev_size[1] + 2.f, (float)g.bev_size[2] + 2.f);
        const uint64_t tiles = (uint64_t)pa.tx * pa.ty * pa.tz;
        if (tiles > 0xffffffffULL) return set_error(CCCS_EGEOM, "PVCS box too large");
        const double nbox = (double)pa.sx * pa.sy * pa.sz;
        // cones in passes: the REV data the resident tiles gather at a time then belongs to fewer
        // cones (better L2 / Infinity Cache reuse); each pass adds its cones' dose to the volume.
        // Automatic (0): as many cones per pass as keep a pass's REV data within 128 MB (measured
        // on the 256x256x180 serving requests, ~31 MB per cone: 4 cones per pass -20 %; small plans
        // stay in one pass, where extra passes only add launches)
        int per = ctx->sched.pvcs_pass_cones;
        if (per <= 0) {
            const double cone_bytes = 4.0 * (double)stored / std::max(1, g.nrays);
            per = (int)std::min<double>(g.nrays, std::max(1.0, std::floor(128.0 * 1048576.0 / std::max(cone_bytes, 1.0))));
        }
        int tk = prof_begin(ctx, K_PVCS);
        for (pa.q0 = 0; pa.q0 < g.nrays; pa.q0 += per) {
            pa.q1 = std::min(g.nrays, pa.q0 + per);
            hipLaunchKernelGGL(pvcs_kernel, dim3((unsigned)tiles), dim3(256), 0, ctx->stream, pa);
        }
        // algorithmic traffic: every stored sample read once, the box's dose read and written once
        prof_end(ctx, tk, 4.0 * (double)stored + 8.0 * nbox, 0.0, (uint64_t)nbox);
        CCCS_HIP(hipGetLastError());
    }
    return CCCS_OK;
}

int upload_fmap(cccs_ctx* ctx, const cccs_beam& b, const float* host_fmap, const float** d_out) {
    size_t n = (size_t)b.fmap_size[0] * b.fmap_size[1];
    CCCS_TRY(ctx->fmap.ensure(n * sizeof(float)));
    // synchronous copy: the host buffer may be reused by the caller right after the call
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_HIP(hipMemcpy(ctx->fmap.p, host_fmap, n * sizeof(float), hipMemcpyHostToDevice));
    *d_out = ctx->fmap.as<float>();
    return CCCS_OK;
}

}  // namespace

int beam_dose_scratch(BeamScratch& S, const cccs_beam& b, const float* d_fmap, float* d_dose) {
    cccs_ctx* ctx = S.ctx;
    CCCS_TRY(check_beam(ctx, b));
    float* d_terma = nullptr;
    CCCS_TRY(S.get(ctx->terma, bbox_voxels(ctx), &d_terma));
    TermaArgs ta = terma_args(ctx, b, d_fmap, d_terma, -1);
    if (const float* unit = angle_cache_lookup(ctx, b)) CCCS_TRY(launch_terma_from_unit(ctx, ta, unit));
    else CCCS_TRY(launch_terma(ctx, ta));
    return run_cones(S, b, d_terma, d_dose, nullptr);
}

// Shared by the C-ABI entry points below and by the beamlet path.
int beam_dose_one(cccs_ctx* ctx, const cccs_beam& b, const float* d_fmap, float* d_dose) {
    BeamScratch S;
    CCCS_TRY(context_scratch(ctx, S));
    return beam_dose_scratch(S, b, d_fmap, d_dose);
}

}  // namespace cccs

// ---- angle cache: resident unit TERMA per beam source -------------------------------------------
// The unit-fluence TERMA of a full-beam dose depends on the beam's source and SAD (the Siddon
// march and the inverse square) and on the problem (density, spectrum, grids), not on the fluence
// map or the collimator.  An optimiser asks for beams from a fixed candidate set, so each
// candidate source's unit TERMA over the calc_bbox can stay in HBM; a request's TERMA is then the
// voxel's fluence times that entry, bit-identical to the march (the direct path multiplies in the
// same order).
struct cccs_angle_cache {
    int device = 0;
    uint32_t size[3] = {0, 0, 0}, bbox_start[3] = {0, 0, 0}, bbox_size[3] = {0, 0, 0};
    float voxel[3] = {0, 0, 0}, start[3] = {0, 0, 0};
    float bhc = 1.f;
    uint64_t dens_hash = 0;
    cccs::SpecDev spec{};
    size_t nbbox = 0;
    float* data = nullptr;  // one calc_bbox grid per entry
    std::map<std::array<uint32_t, 4>, size_t> index;  // (source bits, SAD bits) -> entry
};

namespace cccs {
namespace {
std::array<uint32_t, 4> angle_key(const cccs_beam& b) {
    std::array<uint32_t, 4> k;
    std::memcpy(&k[0], &b.source[0], 4);
    std::memcpy(&k[1], &b.source[1], 4);
    std::memcpy(&k[2], &b.source[2], 4);
    std::memcpy(&k[3], &b.sad, 4);
    return k;
}

bool cache_fits(const cccs_ctx* ctx, const cccs_angle_cache* c) {
    const cccs_problem& p = ctx->prob;
    for (int k = 0; k < 3; k++)
        if (c->size[k] != p.size[k] || c->bbox_start[k] != p.bbox_start[k] || c->bbox_size[k] != p.bbox_size[k] ||
            c->voxel[k] != p.voxel[k] || c->start[k] != p.start[k])
            return false;
    if (c->bhc != p.beamhard_correct || c->dens_hash != ctx->dens_hash || c->device != ctx->device) return false;
    return std::memcmp(&c->spec, &ctx->spec, sizeof(SpecDev)) == 0;
}

const float* angle_cache_lookup(const cccs_ctx* ctx, const cccs_beam& b) {
    const cccs_angle_cache* c = ctx->angle_cache;
    if (!c || ctx->prob.ss_factor > 1) return nullptr;
    auto it = c->index.find(angle_key(b));
    return it == c->index.end() ? nullptr : c->data + it->second * c->nbbox;
}
}  // namespace

// ---- fixed scratch of the no-allocation entry point ----------------------------------------------
// The full-beam path's device scratch (TERMA, interleaved density/TERMA, line table, scan partials,
// cone constants, compact REV store) for callers that cannot allocate: one array per GPU, held
// until the process ends.  Host code uses the per-device address from hipGetSymbolAddress, never
// the array's name.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return set_error(CCCS_EINVAL, "device ordinal beyond 64");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return CCCS_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_error(CCCS_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return CCCS_OK;
}

}  // namespace cccs

using namespace cccs;

// =================================================================================================
// C-ABI

extern "C" int cccs_angle_cache_create(cccs_ctx* ctx, const cccs_beam* beams, int nbeams, uint64_t max_bytes,
                                      cccs_angle_cache** out) {
    clear_error();
    if (!ctx || !out || (!beams && nbeams > 0) || nbeams < 0) return set_error(CCCS_EINVAL, "null argument");
    *out = nullptr;
    if (ctx->prob.ss_factor > 1) return set_error(CCCS_EINVAL, "the angle cache needs ss_factor 1");
    CCCS_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<cccs_angle_cache> c(new cccs_angle_cache());
    const cccs_problem& p = ctx->prob;
    for (int k = 0; k < 3; k++) {
        c->size[k] = p.size[k];
        c->bbox_start[k] = p.bbox_start[k];
        c->bbox_size[k] = p.bbox_size[k];
        c->voxel[k] = p.voxel[k];
        c->start[k] = p.start[k];
    }
    c->bhc = p.beamhard_correct;
    c->dens_hash = ctx->dens_hash;
    c->spec = ctx->spec;
    c->device = ctx->device;
    c->nbbox = bbox_voxels(ctx);
    const uint64_t per = sizeof(float) * (uint64_t)c->nbbox;
    std::vector<int> keep;
    for (int i = 0; i < nbeams; i++) {
        // only the source and SAD enter the unit TERMA; the fluence-map fields are not used
        const cccs_beam& b = beams[i];
        if (!(b.sad > 0.f) || !std::isfinite(b.source[0]) || !std::isfinite(b.source[1]) || !std::isfinite(b.source[2]))
            return set_error(CCCS_EINVAL, "angle cache: beam " + std::to_string(i) + " needs a finite source and sad > 0");
        const auto k = angle_key(beams[i]);
        if (c->index.count(k)) continue;
        if ((uint64_t)(keep.size() + 1) * per > max_bytes) break;
        c->index[k] = keep.size();
        keep.push_back(i);
    }
    if (!keep.empty()) {
        if (hipMalloc(&c->data, per * keep.size()) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(CCCS_ENOMEM, "angle cache of " + std::to_string(per * keep.size()) + " bytes");
        }
        for (size_t e = 0; e < keep.size(); e++) {
            TermaArgs a = terma_args(ctx, beams[keep[e]], nullptr, nullptr, -1);
            dim3 blk(64, 4, 1), grd((a.bx + 63) / 64, (a.by + 3) / 4, a.bz);
            hipLaunchKernelGGL(unit_terma_kernel, grd, blk, 0, ctx->stream, a, c->data + e * c->nbbox);
            const hipError_t err = hipGetLastError();
            if (err != hipSuccess) {
                (void)hipStreamSynchronize(ctx->stream);
                (void)hipFree(c->data);
                return set_error(CCCS_EDEVICE, std::string("unit TERMA kernel: ") + hipGetErrorString(err));
            }
        }
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
            (void)hipFree(c->data);
            return set_error(CCCS_EDEVICE, "angle cache build");
        }
    }
    *out = c.release();
    return CCCS_OK;
}

extern "C" int cccs_angle_cache_info(const cccs_angle_cache* cache, int* nangles, uint64_t* bytes) {
    clear_error();
    if (!cache) return set_error(CCCS_EINVAL, "null argument");
    if (nangles) *nangles = (int)cache->index.size();
    if (bytes) *bytes = sizeof(float) * (uint64_t)cache->nbbox * cache->index.size();
    return CCCS_OK;
}

extern "C" int cccs_ctx_use_angle_cache(cccs_ctx* ctx, const cccs_angle_cache* cache) {
    clear_error();
    if (!ctx) return set_error(CCCS_EINVAL, "null argument");
    if (cache && !cache_fits(ctx, cache))
        return set_error(CCCS_EINVAL, "angle cache built for another problem, density or device");
    ctx->angle_cache = cache;
    return CCCS_OK;
}

extern "C" int cccs_angle_cache_destroy(cccs_angle_cache* cache) {
    clear_error();
    if (!cache) return CCCS_OK;
    if (cache->data) {
        CCCS_HIP(hipSetDevice(cache->device));
        CCCS_HIP(hipFree(cache->data));
    }
    delete cache;
    return CCCS_OK;
}

extern "C" int cccs_beam_dose_device_noalloc(cccs_ctx* ctx, const cccs_beam* beams, int nbeams,
                                             const float* d_fluence, float* d_dose) {
    clear_error();
    if (!ctx || (!beams && nbeams > 0) || nbeams < 0 || !d_dose || (!d_fluence && nbeams > 0))
        return set_error(CCCS_EINVAL, "null argument");
    if (ctx->prof.on) return set_error(CCCS_EINVAL, "per-kernel profiling allocates: turn it off for this call");
    CCCS_HIP(hipSetDevice(ctx->device));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch_ptr[ctx->device];
    }
    if (!base) return set_error(CCCS_EDEVICE, "fixed scratch not resolved for this device");
    float host_small[16];
    ConeDev host_cones[kMaxRays];
    size_t fo = 0;
    for (int i = 0; i < nbeams; i++) {
        BeamScratch S;
        S.ctx = ctx;
        S.arena = static_cast<char*>(base);
        S.cap = kNoAllocScratchBytes;
        S.host_small = host_small;
        S.host_cones = host_cones;
        CCCS_TRY(beam_dose_scratch(S, beams[i], d_fluence + fo, d_dose));
        fo += (size_t)beams[i].fmap_size[0] * beams[i].fmap_size[1];
    }
    return CCCS_OK;
}

extern "C" int cccs_beam_dose_device(cccs_ctx* ctx, const cccs_beam* beams, int nbeams, const float* d_fluence,
                                     float* d_dose) {
    clear_error();
    if (!ctx || (!beams && nbeams > 0) || nbeams < 0 || !d_dose || (!d_fluence && nbeams > 0))
        return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    size_t fo = 0;
    for (int i = 0; i < nbeams; i++) {
        CCCS_TRY(beam_dose_one(ctx, beams[i], d_fluence + fo, d_dose));
        fo += (size_t)beams[i].fmap_size[0] * beams[i].fmap_size[1];
    }
    return CCCS_OK;
}

extern "C" int cccs_beam_dose(cccs_ctx* ctx, const cccs_beam* beams, int nbeams, const float* const* fluence,
                              float* dose) {
    clear_error();
    if (!ctx || (!beams && nbeams > 0) || nbeams < 0 || !dose || (!fluence && nbeams > 0))
        return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    size_t nv = vol_voxels(ctx);
    CCCS_TRY(ctx->dose.ensure(nv * sizeof(float)));
    float* d = ctx->dose.as<float>();
    CCCS_HIP(hipMemsetAsync(d, 0, nv * sizeof(float), ctx->stream));
    for (int i = 0; i < nbeams; i++) {
        if (!fluence[i]) return set_error(CCCS_EINVAL, "null fluence map");
        const float* dfm = nullptr;
        CCCS_TRY(upload_fmap(ctx, beams[i], fluence[i], &dfm));
        CCCS_TRY(beam_dose_one(ctx, beams[i], dfm, d));
    }
    return download_staged(ctx, dose, d, nv * sizeof(float), true);
}

extern "C" int cccs_beam_terma(cccs_ctx* ctx, const cccs_beam* beam, const float* fluence, int beamletnum,
                               float* terma_bbox) {
    clear_error();
    if (!ctx || !beam || !fluence || !terma_bbox) return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    CCCS_TRY(check_beam(ctx, *beam));
    const float* dfm = nullptr;
    CCCS_TRY(upload_fmap(ctx, *beam, fluence, &dfm));
    size_t n = bbox_voxels(ctx);
    CCCS_TRY(ctx->terma.ensure(n * sizeof(float)));
    TermaArgs ta = terma_args(ctx, *beam, dfm, ctx->terma.as<float>(), beamletnum);
    CCCS_TRY(launch_terma(ctx, ta));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_HIP(hipMemcpy(terma_bbox, ctx->terma.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return CCCS_OK;
}

extern "C" int cccs_beam_geometry_of(cccs_ctx* ctx, const cccs_beam* beam, const float* fluence,
                                     cccs_beam_geometry* geom) {
    clear_error();
    if (!ctx || !beam || !fluence || !geom) return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    CCCS_TRY(check_beam(ctx, *beam));
    const float* dfm = nullptr;
    CCCS_TRY(upload_fmap(ctx, *beam, fluence, &dfm));
    CCCS_TRY(ctx->terma.ensure(sizeof(float) * bbox_voxels(ctx)));
    TermaArgs ta = terma_args(ctx, *beam, dfm, ctx->terma.as<float>(), -1);
    CCCS_TRY(launch_terma(ctx, ta));
    BeamGeom g;
    BeamScratch S;
    CCCS_TRY(context_scratch(ctx, S));
    CCCS_TRY(run_cones(S, *beam, ctx->terma.as<float>(), nullptr, &g));
    std::memset(geom, 0, sizeof(*geom));
    geom->lim_min[0] = g.lim_min.x; geom->lim_min[1] = g.lim_min.y; geom->lim_min[2] = g.lim_min.z;
    for (int k = 0; k < 3; k++) geom->bev_size[k] = g.bev_size[k];
    geom->bev_min[0] = g.bev_min.x; geom->bev_min[1] = g.bev_min.y; geom->bev_min[2] = g.bev_min.z;
    geom->bev_max[0] = g.bev_max.x; geom->bev_max[1] = g.bev_max.y; geom->bev_max[2] = g.bev_max.z;
    geom->nrays = g.nrays;
    for (int r = 0; r < g.nrays; r++) {
        geom->rev_min[r][0] = g.rev_min[r].x; geom->rev_min[r][1] = g.rev_min[r].y; geom->rev_min[r][2] = g.rev_min[r].z;
        for (int k = 0; k < 3; k++) geom->rev_size[r][k] = g.rev_size[r][k];
    }
    return CCCS_OK;
}

extern "C" int cccs_fluence_projection(cccs_ctx* ctx, const uint8_t* mask, const cccs_beam* beams, int nbeams,
                                       float* fluence) {
    clear_error();
    if (!ctx || !mask || (!beams && nbeams > 0) || nbeams < 0 || (!fluence && nbeams > 0))
        return set_error(CCCS_EINVAL, "null argument");
    if (nbeams == 0) return CCCS_OK;
    CCCS_HIP(hipSetDevice(ctx->device));
    std::vector<ProjBeam> pb((size_t)nbeams);
    uint64_t total = 0;
    for (int i = 0; i < nbeams; i++) {
        const cccs_beam& b = beams[i];
        if (b.fmap_size[0] == 0 || b.fmap_size[1] == 0) return set_error(CCCS_EINVAL, "empty fluence map");
        if (!(b.beamlet_size[0] > 0.f) || !(b.beamlet_size[1] > 0.f)) return set_error(CCCS_EINVAL, "beamlet size");
        ProjBeam& o = pb[(size_t)i];
        o.src = ld3(b.source);
        o.iso = ld3(b.isocenter);
        o.inv = rot2_beam_inv(b.azimuth, b.zenith, b.coll);
        o.fx = (int)b.fmap_size[0];
        o.fy = (int)b.fmap_size[1];
        o.bsx = b.beamlet_size[0];
        o.bsy = b.beamlet_size[1];
        o.off = (uint32_t)total;
        total += (uint64_t)b.fmap_size[0] * b.fmap_size[1];
        if (total > 0x7fffffffULL) return set_error(CCCS_EINVAL, "too many beamlets");
    }
    const size_t nv = vol_voxels(ctx);
    // scratch: the mask (bytes, then float), the beam table and the maps; released on every return
    DevBuf mb, mf, tb, ob;
    struct Release {
        DevBuf* b[4];
        ~Release() {
            for (DevBuf* x : b) x->release();
        }
    } release{{&mb, &mf, &tb, &ob}};
    CCCS_TRY(mb.ensure(nv));
    CCCS_TRY(mf.ensure(nv * sizeof(float)));
    CCCS_TRY(tb.ensure(sizeof(ProjBeam) * pb.size()));
    CCCS_TRY(ob.ensure(sizeof(float) * total));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_HIP(hipMemcpy(mb.p, mask, nv, hipMemcpyHostToDevice));
    CCCS_HIP(hipMemcpy(tb.p, pb.data(), sizeof(ProjBeam) * pb.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mask_to_float_kernel, dim3(1024), dim3(256), 0, ctx->stream, (const uint8_t*)mb.p,
                       mf.as<float>(), nv);
    CCCS_HIP(hipGetLastError());
    hipLaunchKernelGGL(projection_kernel, dim3((unsigned)((total + 63) / 64)), dim3(64), 0, ctx->stream,
                       (const float*)mf.as<float>(), ctx->vol, (const ProjBeam*)tb.as<ProjBeam>(), nbeams,
                       (uint32_t)total, ob.as<float>());
    CCCS_HIP(hipGetLastError());
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_HIP(hipMemcpy(fluence, ob.p, sizeof(float) * total, hipMemcpyDeviceToHost));
    return CCCS_OK;
}

extern "C" int cccs_fluence_convolve(cccs_ctx* ctx, const float* fluence, int nmaps, int width, int height,
                                     const float* kernel, int kw, int kh, float* out) {
    clear_error();
    if (!ctx || (nmaps > 0 && (!fluence || !kernel || !out))) return set_error(CCCS_EINVAL, "null argument");
    if (nmaps < 0 || width < 1 || height < 1 || width > 4096 || height > 4096) return set_error(CCCS_EINVAL, "map size");
    if (kw < 1 || kh < 1 || kw > kConvMaxK || kh > kConvMaxK || !(kw & 1) || !(kh & 1))
        return set_error(CCCS_EINVAL, "kernel dims must be odd and at most 31");
    if (nmaps == 0) return CCCS_OK;
    if (nmaps > 65535) return set_error(CCCS_EINVAL, "at most 65535 maps per call");
    CCCS_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)nmaps * width * height;
    DevBuf fb, kb, ob;
    struct Release {
        DevBuf* b[3];
        ~Release() {
            for (DevBuf* x : b) x->release();
        }
    } release{{&fb, &kb, &ob}};
    CCCS_TRY(fb.ensure(sizeof(float) * n));
    CCCS_TRY(kb.ensure(sizeof(float) * (size_t)kw * kh));
    CCCS_TRY(ob.ensure(sizeof(float) * n));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_HIP(hipMemcpy(fb.p, fluence, sizeof(float) * n, hipMemcpyHostToDevice));
    CCCS_HIP(hipMemcpy(kb.p, kernel, sizeof(float) * (size_t)kw * kh, hipMemcpyHostToDevice));
    const dim3 grid((unsigned)((width + kConvTile - 1) / kConvTile), (unsigned)((height + kConvTile - 1) / kConvTile),
                    (unsigned)nmaps);
    hipLaunchKernelGGL(fluence_conv_kernel, grid, dim3(256), 0, ctx->stream, (const float*)fb.as<float>(), width, height,
                       (const float*)kb.as<float>(), kw, kh, ob.as<float>());
    CCCS_HIP(hipGetLastError());
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_HIP(hipMemcpy(out, ob.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return CCCS_OK;
}

extern "C" int cccs_siddon_paths(cccs_ctx* ctx, const float* source, const float* points, int npoints,
                                 int max_segments, int32_t* ijk, float* mid, float* len, int32_t* counts, float* rpl) {
    clear_error();
    if (!ctx || !source || !points || npoints < 0 || max_segments < 0 || !ijk || !mid || !len || !counts || !rpl)
        return set_error(CCCS_EINVAL, "null argument");
    if (npoints == 0) return CCCS_OK;
    CCCS_HIP(hipSetDevice(ctx->device));
    const size_t ns = (size_t)npoints * std::max(1, max_segments);
    DevBuf b;
    size_t bytes = sizeof(float) * 3 * npoints + ns * (3 * sizeof(int32_t) + 3 * sizeof(float) + sizeof(float)) +
                   (size_t)npoints * (sizeof(int32_t) + sizeof(float));
    CCCS_TRY(b.ensure(bytes));
    char* base = b.as<char>();
    float* d_pts = (float*)base;
    int32_t* d_ijk = (int32_t*)(d_pts + 3 * npoints);
    float* d_mid = (float*)(d_ijk + 3 * ns);
    float* d_len = d_mid + 3 * ns;
    int32_t* d_cnt = (int32_t*)(d_len + ns);
    float* d_rpl = (float*)(d_cnt + npoints);
    int st = CCCS_OK;
    do {
        if (hipMemcpy(d_pts, points, sizeof(float) * 3 * npoints, hipMemcpyHostToDevice) != hipSuccess) { st = set_error(CCCS_EDEVICE, "copy"); break; }
        // slots past a point's segment count are returned as zeros
        if (hipMemset(d_ijk, 0, bytes - sizeof(float) * 3 * npoints) != hipSuccess) { st = set_error(CCCS_EDEVICE, "memset"); break; }
        PathArgs a;
        a.dens = ctx->d_dens;
        a.vd = ctx->vol;
        a.src = ld3(source);
        a.pts = d_pts;
        a.n = npoints;
        a.maxseg = max_segments;
        a.ijk = d_ijk; a.mid = d_mid; a.len = d_len; a.counts = d_cnt; a.rpl = d_rpl;
        hipLaunchKernelGGL(siddon_path_kernel, dim3((npoints + 255) / 256), dim3(256), 0, ctx->stream, a);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) { st = set_error(CCCS_EDEVICE, "siddon kernel"); break; }
        if (max_segments > 0) {
            if (hipMemcpy(ijk, d_ijk, ns * 3 * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(mid, d_mid, ns * 3 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(len, d_len, ns * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) { st = set_error(CCCS_EDEVICE, "copy back"); break; }
        }
        if (hipMemcpy(counts, d_cnt, npoints * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(rpl, d_rpl, npoints * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) { st = set_error(CCCS_EDEVICE, "copy back"); break; }
    } while (false);
    b.release();
    return st;
}

