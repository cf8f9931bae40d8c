// cccs_internal.hpp -- types and math shared by the host pipeline and the gfx950 kernels.
//
// Geometry convention (CCCS/CudaUtilities/geometry.cu:6-19): PVCS is the patient volume frame
// (x fastest in memory, `start` = centre of voxel 0); BEV is the beam frame with the beam along
// +y; REV is the frame of one convolution (cone) direction, sampled as lines along its y axis.
// Rotations are applied as 3x3 matrices built once per beam/cone on the host in double
// precision and rounded to FP32, instead of per-thread Rodrigues evaluations.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/cccs.h"

#define CCCS_HD __host__ __device__ __forceinline__

namespace cccs {

struct v3 {
    float x, y, z;
};
CCCS_HD v3 mk3(float x, float y, float z) { return v3{x, y, z}; }
CCCS_HD v3 operator+(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
CCCS_HD v3 operator-(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
CCCS_HD v3 operator*(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
CCCS_HD v3 operator/(v3 a, v3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
CCCS_HD v3 operator*(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
CCCS_HD v3 operator*(float s, v3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
CCCS_HD v3 operator+(v3 a, float s) { return mk3(a.x + s, a.y + s, a.z + s); }
CCCS_HD v3 operator-(v3 a, float s) { return mk3(a.x - s, a.y - s, a.z - s); }
CCCS_HD float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CCCS_HD float len3(v3 v) { return sqrtf(dot3(v, v)); }
CCCS_HD v3 norm3(v3 v) {
    float inv = 1.0f / sqrtf(dot3(v, v));
    return v * inv;
}
inline v3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }
inline v3 ldu3(const uint32_t* p) { return mk3((float)p[0], (float)p[1], (float)p[2]); }

// Row-major 3x3, FP32.
struct m33 {
    float m[9];
    CCCS_HD v3 operator()(v3 v) const {
        return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
                   m[6] * v.x + m[7] * v.y + m[8] * v.z);
    }
};

// Two successive axis rotations in the Rodrigues form of CudaUtilities/geometry.cu:46-58, with the
// sines/cosines computed once on the host (libm sincosf).  Used where a rounding difference would
// flip a discrete decision (fluence-cell lookup, crop-box size, REV grid size), so those decisions
// are reproducible against a host evaluation of the same formula.
struct Rot2 {
    v3 a1;
    float s1, c1;
    v3 a2;
    float s2, c2;
};
CCCS_HD v3 rodrigues(v3 p, v3 r, float s, float c) {
    return mk3((-r.x * (-r.x * p.x - r.y * p.y - r.z * p.z)) * (1 - c) + p.x * c + (-r.z * p.y + r.y * p.z) * s,
               (-r.y * (-r.x * p.x - r.y * p.y - r.z * p.z)) * (1 - c) + p.y * c + (+r.z * p.x - r.x * p.z) * s,
               (-r.z * (-r.x * p.x - r.y * p.y - r.z * p.z)) * (1 - c) + p.z * c + (-r.y * p.x + r.x * p.y) * s);
}
CCCS_HD v3 rot2(const Rot2& R, v3 p) { return rodrigues(rodrigues(p, R.a1, R.s1, R.c1), R.a2, R.s2, R.c2); }
// host builders (CudaUtilities/geometry.cu:69-75, 85-92, 126-131, 140-145)
Rot2 rot2_beam_fwd(float theta, float phi, float coll);
Rot2 rot2_beam_inv(float theta, float phi, float coll);
Rot2 rot2_kernel_fwd(float theta, float phi);
Rot2 rot2_kernel_inv(float theta, float phi);

// float -> int toward -inf with NaN -> 0 (the convention of the reference's float->int casts)
CCCS_HD int f2i_rd(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483520.f) return 2147483647;
    if (x <= -2147483648.f) return (-2147483647 - 1);
    return (int)floorf(x);
}

// ---- host: rotation matrices (double precision build) ------------------------------------------
struct dm33 {
    double m[9];
};
dm33 dm_axis(double ax, double ay, double az, double t);  // right-hand rotation by t about unit axis
dm33 dm_mul(const dm33& a, const dm33& b);
dm33 dm_transpose(const dm33& a);
m33 to_f(const dm33& a);
v3 dm_apply_f(const dm33& a, v3 v);  // apply in double, round result to float

// BEV <- PVCS and inverse (CudaUtilities/geometry.cu:61-92)
dm33 dm_beam_fwd(float theta, float phi, float coll);
dm33 dm_beam_inv(float theta, float phi, float coll);
// REV <- BEV and inverse (geometry.cu:119-145)
dm33 dm_kernel_fwd(float theta, float phi);
dm33 dm_kernel_inv(float theta, float phi);
// beamlet bev <- BEV and inverse (geometry.cu:95-115)
dm33 dm_beamlet_fwd(float theta, float phi);
dm33 dm_beamlet_inv(float theta, float phi);

// ---- trilinear sampling with texture-coordinate conventions ------------------------------------
// Texel i is centred at coordinate i+0.5; taps outside [0,n) read 0 (border addressing).
CCCS_HD float fetch_or0(const float* __restrict__ v, int nx, int ny, int nz, int i, int j, int k) {
    return (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz)
               ? 0.f
               : v[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)];
}
CCCS_HD float tex3_border(const float* __restrict__ v, int nx, int ny, int nz, float u, float w, float t) {
    float xb = u - 0.5f, yb = w - 0.5f, zb = t - 0.5f;
    if (!(xb > -1.f && yb > -1.f && zb > -1.f && xb < (float)nx && yb < (float)ny && zb < (float)nz)) return 0.f;
    float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    int i = (int)fx, j = (int)fy, k = (int)fz;
    float a = xb - fx, b = yb - fy, c = zb - fz;
    if (i >= 0 && j >= 0 && k >= 0 && i + 1 < nx && j + 1 < ny && k + 1 < nz) {
        // all eight taps inside: the same interpolation without per-tap bounds tests
        const size_t sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
        const float* p = v + (size_t)i + sy * (size_t)j + sz * (size_t)k;
        float e0 = p[0] + a * (p[1] - p[0]), e1 = p[sy] + a * (p[sy + 1] - p[sy]);
        float f0 = p[sz] + a * (p[sz + 1] - p[sz]), f1 = p[sz + sy] + a * (p[sz + sy + 1] - p[sz + sy]);
        float g0 = e0 + b * (e1 - e0), g1 = f0 + b * (f1 - f0);
        return g0 + c * (g1 - g0);
    }
    float c00 = fetch_or0(v, nx, ny, nz, i, j, k), c10 = fetch_or0(v, nx, ny, nz, i + 1, j, k);
    float c01 = fetch_or0(v, nx, ny, nz, i, j + 1, k), c11 = fetch_or0(v, nx, ny, nz, i + 1, j + 1, k);
    float d00 = fetch_or0(v, nx, ny, nz, i, j, k + 1), d10 = fetch_or0(v, nx, ny, nz, i + 1, j, k + 1);
    float d01 = fetch_or0(v, nx, ny, nz, i, j + 1, k + 1), d11 = fetch_or0(v, nx, ny, nz, i + 1, j + 1, k + 1);
    float e0 = c00 + a * (c10 - c00), e1 = c01 + a * (c11 - c01);
    float f0 = d00 + a * (d10 - d00), f1 = d01 + a * (d11 - d01);
    float g0 = e0 + b * (e1 - e0), g1 = f0 + b * (f1 - f0);
    return g0 + c * (g1 - g0);
}

// tex3_border over an interleaved {density, TERMA} volume: both channels at one point with one set
// of taps (8-byte loads, one address per tap).  Each channel follows tex3_border's operation
// sequence, so a channel equals tex3_border of that channel's volume at the same coordinates.
CCCS_HD float2 fetch2_or0(const float2* __restrict__ v, int nx, int ny, int nz, int i, int j, int k) {
    return (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz)
               ? make_float2(0.f, 0.f)
               : v[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)];
}
CCCS_HD float lerp3(float c00, float c10, float c01, float c11, float d00, float d10, float d01, float d11, float a,
                    float b, float c) {
    float e0 = c00 + a * (c10 - c00), e1 = c01 + a * (c11 - c01);
    float f0 = d00 + a * (d10 - d00), f1 = d01 + a * (d11 - d01);
    float g0 = e0 + b * (e1 - e0), g1 = f0 + b * (f1 - f0);
    return g0 + c * (g1 - g0);
}
CCCS_HD float2 tex3_pair(const float2* __restrict__ v, int nx, int ny, int nz, float u, float w, float t) {
    float xb = u - 0.5f, yb = w - 0.5f, zb = t - 0.5f;
    if (!(xb > -1.f && yb > -1.f && zb > -1.f && xb < (float)nx && yb < (float)ny && zb < (float)nz))
        return make_float2(0.f, 0.f);
    float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    int i = (int)fx, j = (int)fy, k = (int)fz;
    float a = xb - fx, b = yb - fy, c = zb - fz;
    float2 c00, c10, c01, c11, d00, d10, d01, d11;
    if (i >= 0 && j >= 0 && k >= 0 && i + 1 < nx && j + 1 < ny && k + 1 < nz) {
        // volumes below 2^31 voxels (checked by the caller): 32-bit tap offsets
        const uint32_t sy = (uint32_t)nx, sz = (uint32_t)nx * (uint32_t)ny;
        const float2* p = v + ((uint32_t)i + sy * (uint32_t)j + sz * (uint32_t)k);
        c00 = p[0]; c10 = p[1]; c01 = p[sy]; c11 = p[sy + 1];
        d00 = p[sz]; d10 = p[sz + 1]; d01 = p[sz + sy]; d11 = p[sz + sy + 1];
    } else {
        c00 = fetch2_or0(v, nx, ny, nz, i, j, k); c10 = fetch2_or0(v, nx, ny, nz, i + 1, j, k);
        c01 = fetch2_or0(v, nx, ny, nz, i, j + 1, k); c11 = fetch2_or0(v, nx, ny, nz, i + 1, j + 1, k);
        d00 = fetch2_or0(v, nx, ny, nz, i, j, k + 1); d10 = fetch2_or0(v, nx, ny, nz, i + 1, j, k + 1);
        d01 = fetch2_or0(v, nx, ny, nz, i, j + 1, k + 1); d11 = fetch2_or0(v, nx, ny, nz, i + 1, j + 1, k + 1);
    }
    return make_float2(lerp3(c00.x, c10.x, c01.x, c11.x, d00.x, d10.x, d01.x, d11.x, a, b, c),
                       lerp3(c00.y, c10.y, c01.y, c11.y, d00.y, d10.y, d01.y, d11.y, a, b, c));
}

// ---- Siddon march (DoseCalcAlgo/cudaSiddon.cu:20-115) -------------------------------------------
// The float operation sequence is fixed (no contraction: every file is built with
// -ffp-contract=off) so the segment list is reproducible bit for bit on any IEEE FP32 machine.
// `seg(fx, fy, fz, len)` is called for every segment of length >= 0.001 cm with the midpoint
// fetch coordinate (density is sampled at that coordinate without the +0.5 texel shift).
template <class Seg>
CCCS_HD float siddon_march(v3 this_voxel, v3 start, v3 end, v3 voxel, v3 src, const float* __restrict__ dens,
                           int nx, int ny, int nz, Seg&& seg, int* iters) {
    v3 diff = this_voxel - src;
    v3 startdiff = start - src;
    v3 enddiff = end - src;
    float ax0 = ((startdiff.x) - 0.5f * voxel.x) / (diff.x);
    float ay0 = ((startdiff.y) - 0.5f * voxel.y) / (diff.y);
    float az0 = ((startdiff.z) - 0.5f * voxel.z) / (diff.z);
    float ax1 = ((enddiff.x) + 0.5f * voxel.x) / (diff.x);
    float ay1 = ((enddiff.y) + 0.5f * voxel.y) / (diff.y);
    float az1 = ((enddiff.z) + 0.5f * voxel.z) / (diff.z);
    float a_min = fmaxf(0.f, fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), fminf(az0, az1)));
    float a_max = fminf(1.f, fminf(fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)), fmaxf(az0, az1)));
    float rpl = 0.0f;
    int iter = 0;
    if (!(a_min >= a_max)) {
        float d12 = len3(diff);
        float step_x = fabsf(voxel.x / diff.x);
        float step_y = fabsf(voxel.y / diff.y);
        float step_z = fabsf(voxel.z / diff.z);
        float alpha = a_min, alpha_x = a_min, alpha_y = a_min, alpha_z = a_min;
        float nextalpha = 0.f;
        while (alpha < a_max && ++iter < 5000) {
            bool vx = (alpha_x >= 0.0f && alpha_x < 1.0f);
            bool vy = (alpha_y >= 0.0f && alpha_y < 1.0f);
            bool vz = (alpha_z >= 0.0f && alpha_z < 1.0f);
            if (!(vx || vy || vz)) break;
            // the three plane tests run in sequence and each sees the alphas already advanced in
            // this step (cudaSiddon.cu:84-95): ties advance every tied plane, and a crossing that
            // lies between the old and the advanced alpha_x can merge into this segment
            if (vx && (!vy || alpha_x <= alpha_y) && (!vz || alpha_x <= alpha_z)) { nextalpha = alpha_x; alpha_x += step_x; }
            if (vy && (!vx || alpha_y <= alpha_x) && (!vz || alpha_y <= alpha_z)) { nextalpha = alpha_y; alpha_y += step_y; }
            if (vz && (!vx || alpha_z <= alpha_x) && (!vy || alpha_z <= alpha_y)) { nextalpha = alpha_z; alpha_z += step_z; }
            float inter = fabsf(d12 * (nextalpha - alpha));
            if (inter >= 0.001f) {
                float am = (nextalpha + alpha) * 0.5f;
                float fx = (src.x + am * diff.x - start.x) / voxel.x;
                float fy = (src.y + am * diff.y - start.y) / voxel.y;
                float fz = (src.z + am * diff.z - start.z) / voxel.z;
                rpl += inter * tex3_border(dens, nx, ny, nz, fx, fy, fz);
                seg(fx, fy, fz, inter);
            }
            alpha = nextalpha;
        }
    }
    if (iters) *iters = iter;
    return rpl;
}

// Radiological path length of the same segment structure as siddon_march (identical plane
// sequence and segment lengths), with the midpoint fetch coordinate formed by multiplying with the
// reciprocal voxel size instead of dividing: the fetch coordinates may differ in the last bit, so
// rpl agrees with siddon_march to FP32 rounding, not bit for bit.  Used for TERMA (both paths).
CCCS_HD float siddon_rpl(v3 this_voxel, v3 start, v3 end, v3 voxel, v3 ivoxel, v3 src, const float* __restrict__ dens,
                         int nx, int ny, int nz, int* iters) {
    v3 diff = this_voxel - src;
    v3 startdiff = start - src;
    v3 enddiff = end - src;
    float ax0 = ((startdiff.x) - 0.5f * voxel.x) / (diff.x);
    float ay0 = ((startdiff.y) - 0.5f * voxel.y) / (diff.y);
    float az0 = ((startdiff.z) - 0.5f * voxel.z) / (diff.z);
    float ax1 = ((enddiff.x) + 0.5f * voxel.x) / (diff.x);
    float ay1 = ((enddiff.y) + 0.5f * voxel.y) / (diff.y);
    float az1 = ((enddiff.z) + 0.5f * voxel.z) / (diff.z);
    float a_min = fmaxf(0.f, fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), fminf(az0, az1)));
    float a_max = fminf(1.f, fminf(fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)), fmaxf(az0, az1)));
    float rpl = 0.0f;
    int iter = 0;
    if (!(a_min >= a_max)) {
        float d12 = len3(diff);
        float step_x = fabsf(voxel.x / diff.x);
        float step_y = fabsf(voxel.y / diff.y);
        float step_z = fabsf(voxel.z / diff.z);
        const v3 o = (src - start) * ivoxel, dd = diff * ivoxel;
        float alpha = a_min, alpha_x = a_min, alpha_y = a_min, alpha_z = a_min;
        float nextalpha = 0.f;
        while (alpha < a_max && ++iter < 5000) {
            bool vx = (alpha_x >= 0.0f && alpha_x < 1.0f);
            bool vy = (alpha_y >= 0.0f && alpha_y < 1.0f);
            bool vz = (alpha_z >= 0.0f && alpha_z < 1.0f);
            if (!(vx || vy || vz)) break;
            if (vx && (!vy || alpha_x <= alpha_y) && (!vz || alpha_x <= alpha_z)) { nextalpha = alpha_x; alpha_x += step_x; }
            if (vy && (!vx || alpha_y <= alpha_x) && (!vz || alpha_y <= alpha_z)) { nextalpha = alpha_y; alpha_y += step_y; }
            if (vz && (!vx || alpha_z <= alpha_x) && (!vy || alpha_z <= alpha_y)) { nextalpha = alpha_z; alpha_z += step_z; }
            float inter = fabsf(d12 * (nextalpha - alpha));
            if (inter >= 0.001f) {
                float am = (nextalpha + alpha) * 0.5f;
                rpl += inter * tex3_border(dens, nx, ny, nz, o.x + am * dd.x, o.y + am * dd.y, o.z + am * dd.z);
            }
            alpha = nextalpha;
        }
    }
    if (iters) *iters = iter;
    return rpl;
}

// ---- device-side problem constants -------------------------------------------------------------
struct VolDesc {
    int nx, ny, nz;
    v3 start, voxel, end;
};

struct BeamDev {
    v3 src, dir_n, iso;
    float sad;
    float iso_dist;  // |iso - src|
    int fmx, fmy;
    float bsx, bsy;
    Rot2 beam_fwd;  // BEV <- PVCS rotation
};

constexpr int kMaxSpec = 30;
struct SpecDev {
    int n;
    float flu[kMaxSpec], en[kMaxSpec], mu[kMaxSpec];
};

// Fluence-map cell seen by point p, -1 outside the map: the index part of the beamlet-cell lookup
// of cudaSiddon_device (cudaSiddon.cu:146-165).
CCCS_HD int fluence_cell(v3 p, const BeamDev& b) {
    v3 diff = p - b.src;
    v3 projection = dot3(diff, b.dir_n) * b.dir_n;
    v3 rejection = diff - projection;
    float trad = len3(rejection) * fabsf(b.iso_dist / len3(projection));
    v3 pf = rot2(b.beam_fwd, trad * norm3(rejection));
    float fx = (pf.x / b.bsx) + ((float)b.fmx * 0.5f);
    float fz = (pf.z / b.bsy) + ((float)b.fmy * 0.5f);
    // a NaN coordinate (voxel on the central axis) passes these tests, as in the reference
    if (fx < 0 || fx >= (float)b.fmx || fz < 0 || fz >= (float)b.fmy) return -1;
    int fidx = f2i_rd(fx) + b.fmx * f2i_rd(fz);
    return (fidx < 0 || fidx >= b.fmx * b.fmy) ? -1 : fidx;
}

// Fluence weight seen by point p (cudaSiddon.cu:146-165): 0 outside the map, for a zero-weight
// cell, or (beamletnum >= 0) for any cell other than beamletnum.
CCCS_HD float point_fluence(v3 p, const float* __restrict__ fmap, const BeamDev& b, int beamletnum) {
    const int fidx = fluence_cell(p, b);
    if (fidx < 0 || (beamletnum >= 0 && fidx != beamletnum)) return 0.f;
    float fluence = fmap[fidx];
    return fluence <= 0.f ? 0.f : fluence;
}

// Spectrum-weighted attenuation sum_e phi_e E_e mu_e exp(-mu_e rpl) for unit fluence
// (cudaSiddon.cu:190-200 with the fluence factored out).
CCCS_HD float unit_terma(float rpl, const SpecDev& sp) {
    float tsum = 0.f;
    for (int e = 0; e < sp.n; e++) tsum += sp.flu[e] * sp.en[e] * sp.mu[e] * expf(-1.f * sp.mu[e] * rpl);
    return tsum;
}

// Inverse-square factor SAD^2 / |p - src|^2 (cudaSiddon.cu:201-204).
CCCS_HD float inv_square(v3 p, const BeamDev& b) {
    float d = len3(p - b.src);
    return (b.sad * b.sad) / (d * d);
}

// TERMA of one point for unit fluence: sum_e phi_e E_e mu_e exp(-mu_e rpl) * bhc * SAD^2/|p - src|^2
// (cudaSiddon.cu:190-204 with the fluence factor taken out).  The full-beam path multiplies it by the
// point's fluence, so a resident copy per beam source (the angle cache) gives its TERMA bit for bit.
CCCS_HD float unit_point_terma(v3 p, const BeamDev& b, const SpecDev& sp, const VolDesc& vd, v3 ivoxel,
                               const float* __restrict__ dens, float bhc, int* iters) {
    const float rpl = siddon_rpl(p, vd.start, vd.end, vd.voxel, ivoxel, b.src, dens, vd.nx, vd.ny, vd.nz, iters);
    return unit_terma(rpl, sp) * bhc * inv_square(p, b);
}

// TERMA of one point (cudaSiddon_device, cudaSiddon.cu:117-208).  Returns 0 outside the field
// or for a zero-weight beamlet; beamletnum >= 0 keeps only that beamlet's footprint.
CCCS_HD float point_terma(v3 p, const float* __restrict__ fmap, const BeamDev& b, const SpecDev& sp,
                          const VolDesc& vd, const float* __restrict__ dens, float bhc, int beamletnum,
                          int* iters) {
    if (iters) *iters = 0;
    const float fluence = point_fluence(p, fmap, b, beamletnum);
    if (fluence <= 0.f) return 0.f;
    float rpl = siddon_march(p, vd.start, vd.end, vd.voxel, b.src, dens, vd.nx, vd.ny, vd.nz,
                             [](float, float, float, float) {}, iters);
    float tsum = 0.f;
    for (int e = 0; e < sp.n; e++) {
        float this_fluence = fluence * sp.flu[e];
        tsum += this_fluence * sp.en[e] * sp.mu[e] * expf(-1.f * sp.mu[e] * rpl);
    }
    return tsum * bhc * inv_square(p, b);
}

// point_terma with the rpl of siddon_rpl (beamlet TERMA).
CCCS_HD float point_terma_rpl(v3 p, const float* __restrict__ fmap, const BeamDev& b, const SpecDev& sp,
                              const VolDesc& vd, v3 ivoxel, const float* __restrict__ dens, float bhc, int beamletnum,
                              int* iters) {
    if (iters) *iters = 0;
    const float fluence = point_fluence(p, fmap, b, beamletnum);
    if (fluence <= 0.f) return 0.f;
    float rpl = siddon_rpl(p, vd.start, vd.end, vd.voxel, ivoxel, b.src, dens, vd.nx, vd.ny, vd.nz, iters);
    float tsum = 0.f;
    for (int e = 0; e < sp.n; e++) {
        float this_fluence = fluence * sp.flu[e];
        tsum += this_fluence * sp.en[e] * sp.mu[e] * expf(-1.f * sp.mu[e] * rpl);
    }
    return tsum * bhc * inv_square(p, b);
}

}  // namespace cccs
