// device_common.hpp -- device helpers shared by the beam and beamlet pipelines: XCD-aware work
// mapping, per-workgroup counters, the REV line table, and the collapsed-cone line walk.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cccs_internal.hpp"

namespace cccs {

// Workgroups are dispatched round-robin over the 8 XCDs of the part (workgroup b runs on XCD
// b % 8).  xcd_remap(b, nb) gives the workgroups of one XCD one contiguous range of work items,
// so neighbouring REV lines / voxel tiles share that XCD's L2.  A bijection on [0, nb).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    const uint32_t nx = 8;
    const uint32_t per = nb / nx, rem = nb % nx;
    const uint32_t x = b % nx, idx = b / nx;
    return x * per + (x < rem ? x : rem) + idx;
}

// Sums one uint32 per thread over the workgroup (LDS tree, <= 256 threads) into part[slot].
// Sums one uint32 per lane over a 64-lane workgroup into part[slot] (no LDS).
__device__ __forceinline__ void wave_count(uint32_t v, uint32_t* __restrict__ part, size_t slot) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (threadIdx.x == 0) part[slot] = v;
}

__device__ __forceinline__ void block_count(uint32_t v, uint32_t* __restrict__ part, size_t slot) {
    __shared__ uint32_t s_cnt[256];
    const int t = threadIdx.x + threadIdx.y * blockDim.x + threadIdx.z * blockDim.x * blockDim.y;
    const int nt = blockDim.x * blockDim.y * blockDim.z;
    s_cnt[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w && t + w < nt) s_cnt[t] += s_cnt[t + w];
        __syncthreads();
    }
    if (t == 0) part[slot] = s_cnt[0];
}

// One REV line of a cone grid in the compact store: samples [i0, i0 + len) of the line are kept at
// rev[off ...]; every other sample of the line is zero.
struct LineMeta {
    uint32_t off;  // first stored sample in the store
    int16_t i0;    // first stored sample index on the line (lines are <= 1024 samples)
    int16_t len;   // stored samples
};
static_assert(sizeof(LineMeta) == 8, "line table entries are 8 bytes");

// Persistent per-line kernels walk lines base, base + G, base + 2G, ... (G = grid size), so that
// neighbouring waves work on neighbouring lines at the same time.  A wave loads the records of
// its next 64 lines at once, one per lane (lane k: line base + k*G), and takes each from its lane:
// one round trip per 64 lines instead of one per line, and lines it skips cost no memory access.
__device__ __forceinline__ LineMeta line_of_lane(const LineMeta& m, int lane) {
    const uint32_t w = (uint32_t)(uint16_t)m.i0 | ((uint32_t)(uint16_t)m.len << 16);
    const uint32_t ow = (uint32_t)__shfl((int)w, lane, 64);
    LineMeta r;
    r.off = (uint32_t)__shfl((int)m.off, lane, 64);
    r.i0 = (int16_t)(ow & 0xffffu);
    r.len = (int16_t)(ow >> 16);
    return r;
}

// Trilinear fetch from a compactly stored REV grid (nx samples per line, ny x nz lines) at texture
// coordinates (u, v, w) = (sample, line y, line z), texel centres at +0.5, border 0: the same
// operation sequence as tex3_border on the dense grid, so both give identical results.
//
// The loads are branch-free and issued in two rounds: the four line records together, then the
// eight samples together, so a lane waits for two memory round trips instead of one per tap.  A
// tap outside the grid reads the grid's first record and a sample outside its line's stored
// interval reads sample 0 of the store (both valid addresses); either is then replaced by zero.
__device__ __forceinline__ float rev_fetch(const LineMeta* __restrict__ meta, const float* __restrict__ rev,
                                           uint64_t line0, int nx, int ny, int nz, float u, float v, float w) {
    if (ny <= 0 || nz <= 0) return 0.f;
    const float ub = u - 0.5f, vb = v - 0.5f, wb = w - 0.5f;
    if (!(ub > -1.f && vb > -1.f && wb > -1.f && ub < (float)nx && vb < (float)ny && wb < (float)nz)) return 0.f;
    const float fu = floorf(ub), fv = floorf(vb), fw = floorf(wb);
    const int i = (int)fu, j = (int)fv, k = (int)fw;
    const float a = ub - fu, b = vb - fv, c = wb - fw;
    LineMeta m[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int jj = j + (q & 1), kk = k + (q >> 1);
        const bool in = jj >= 0 && jj < ny && kk >= 0 && kk < nz;
        m[q] = meta[in ? line0 + (uint64_t)jj + (uint64_t)ny * (uint64_t)kk : line0];
        if (!in) m[q].len = 0;
    }
    float t[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int d = i - m[q].i0;
        const bool vlo = d >= 0 && d < m[q].len, vhi = d + 1 >= 0 && d + 1 < m[q].len;
        const float x0 = rev[vlo ? m[q].off + (uint32_t)d : 0u];
        const float x1 = rev[vhi ? m[q].off + (uint32_t)(d + 1) : 0u];
        const float lo = vlo ? x0 : 0.f, hi = vhi ? x1 : 0.f;
        t[q] = lo + a * (hi - lo);
    }
    const float e0 = t[0] + b * (t[1] - t[0]);
    const float e1 = t[2] + b * (t[3] - t[2]);
    return e0 + c * (e1 - e0);
}

// The line's samples live in LDS as {half step, TERMA}: x = 0.5*delr*density (negative marks a
// sample outside the crop box or the calc_bbox), y = TERMA.
//
// walk_both runs the two walks of destination j (RowConvolve's do/while loops) back to back over
// the line's valid segment [vlo, vhi]: first toward lower indices with the forward kernel row,
// then toward higher indices with the backward row.  Every destination takes exactly vhi - vlo
// steps, so the trip count is the same on every lane of a wave (a scalar loop, no divergence);
// the direction switch and the reference's early exits (radiological depth beyond the last radius,
// reaching the last radial interval) become per-lane flags.  The cumulative kernel is linear
// between two radii, so a lane keeps its current interval in registers and reads the LDS tables
// only when the radiological depth crosses a radius.
CCCS_HD float walk_both(const float2* __restrict__ s, int j, int vlo, int vhi,
                                           const float* __restrict__ R, const float* __restrict__ IW,
                                           const float* __restrict__ rowF, const float* __restrict__ rowB, int nr,
                                           bool active, uint32_t& iters) {
    const float rmax = R[nr - 1];
    const int total = vhi - vlo;
    const int nfwd = j - vlo;
    const float* row = rowF;
    float r = 0.f, last = 0.f, lo = 0.f, hi = R[0], klo = 0.f, slope = rowF[0] * IW[0], sum = 0.f;
    int rs = 0, k = j, dir = -1;
    bool alive = active && nfwd > 0;
    for (int it = 0; it < total; ++it) {
        if (it == nfwd) {  // switch to the backward walk
            dir = 1;
            k = j;
            row = rowB;
            r = 0.f; last = 0.f; lo = 0.f; hi = R[0]; klo = 0.f; slope = rowB[0] * IW[0];
            rs = 0;
            alive = active;
        }
        k += dir;
        const float2 e = s[k];
        r += e.x;
        bool ok = alive && e.x >= 0.f && r <= rmax;
        if (ok && hi < r) {
            do {
                rs++;
                lo = hi;
                hi = R[rs];
                klo = row[rs - 1];
                slope = (row[rs] - klo) * IW[rs];
            } while (hi < r && rs < nr - 1);
        }
        const float cum = klo + (r - lo) * slope;
        sum += ok ? (cum - last) * e.y : 0.f;
        last = cum;
        r += e.x;
        iters += ok ? 1u : 0u;
        alive = ok && rs < nr - 1;
    }
    return sum;
}

// Narrows [t0, t1] to the parameters where o + t*d lies in [lo, hi) (widened by a margin that
// covers FP32 rounding; the exact per-sample test decides membership).
CCCS_HD void clip_slab(float o, float d, float lo, float hi, float& t0, float& t1) {
    if (fabsf(d) < 1e-3f) {
        // the coordinate moves by less than one index unit over a whole line (<= 1024 samples)
        if (!(o >= lo - 2.f && o < hi + 2.f)) { t0 = 1.f; t1 = 0.f; }
        return;
    }
    float ta = (lo - o) / d, tb = (hi - o) / d;
    t0 = fmaxf(t0, fminf(ta, tb));
    t1 = fminf(t1, fmaxf(ta, tb));
}

}  // namespace cccs
