// cccs_oracle.cpp -- CPU oracle for the CCCS dose path.  TEST INFRASTRUCTURE ONLY.
//
// A C++17/OpenMP restatement of the reference algorithm (xuqifan897/EndtoEnd, CCCS/), used as
// the checker for the HIP product and as the host-core baseline of bench.py.  Nothing in the
// product links or includes this file.  It keeps the reference's own geometry (Rodrigues
// rotations with libm sincosf, CudaUtilities/geometry.cu:23-145), its loop structure and its
// quirks; CUDA texture fetches become explicit FP32 trilinear/bilinear interpolation with the
// same coordinate conventions (border 0 for volumes, clamp for the kernel table).
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off).
//
// Pinning (tests/test_oracle_*.py): beam angles by the reference unit test's round trip
// (CCCS/unit_tests/beam.cpp:16-67); kernel tables by aggregation/cumulative-sum invariants;
// Siddon by hand-computed axis-aligned and diagonal paths plus a double-precision brute-force
// line integral; TERMA by the analytic water formula; rotations by inverse/orthonormality;
// the line walk by the telescoping closed form on uniform lines and an O(L^2) double sum;
// resampling by exactness on linear fields; sparsify by hand-built arrays.
#include "cccs_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

// Mutants (test infrastructure): ORC_MUTANT = 1, 2 or 3 builds the oracle with one deliberate
// mistake -- the cone's kernel row taken from its phi index, the beamlet walk on the raw radial
// boundaries, the beamlet unpack half a sample off along the beamlet axis -- so the tests can
// show that their checks catch it (tests/test_oracle_composition.py).
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

namespace {

// float PI as the reference's CUDART_PI_F (include/dosecalc_defs.h:8)
const float kPI = 3.14159265358979323846f;

struct f3 { float x, y, z; };
struct f2 { float x, y; };
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline f3 operator/(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
inline f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
inline f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline f3 operator+(f3 a, float s) { return mk(a.x + s, a.y + s, a.z + s); }
inline f3 operator-(f3 a, float s) { return mk(a.x - s, a.y - s, a.z - s); }
inline f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(f3 v) { return sqrtf(dot(v, v)); }
inline f3 normalize(f3 v) { float inv = 1.0f / sqrtf(dot(v, v)); return v * inv; }
inline f3 fma3(f3 a, f3 b, f3 c) { return mk(fmaf(a.x, b.x, c.x), fmaf(a.y, b.y, c.y), fmaf(a.z, b.z, c.z)); }
inline f3 rcp3(f3 a) { return mk(1.0f / a.x, 1.0f / a.y, 1.0f / a.z); }
inline f3 from_u3(const uint32_t* u) { return mk((float)u[0], (float)u[1], (float)u[2]); }
inline f3 from_f(const float* v) { return mk(v[0], v[1], v[2]); }
inline float comp(const f3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// float -> int rounding toward -inf, NaN -> 0 (the reference's __float2int_rd convention)
inline int f2i_rd(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483520.f) return 2147483647;
    if (x <= -2147483648.f) return (-2147483647 - 1);
    return (int)floorf(x);
}

// ---- Rodrigues rotations, CudaUtilities/geometry.cu:46-145 --------------------------------
inline f3 rot_origin(f3 p, f3 r, float t) {
    float s, c;
    sincosf(t, &s, &c);
    return mk((-r.x * (-r.x * p.x - r.y * p.y - r.z * p.z)) * (1 - c) + p.x * c + (-r.z * p.y + r.y * p.z) * s,
              (-r.y * (-r.x * p.x - r.y * p.y - r.z * p.z)) * (1 - c) + p.y * c + (+r.z * p.x - r.x * p.z) * s,
              (-r.z * (-r.x * p.x - r.y * p.y - r.z * p.z)) * (1 - c) + p.z * c + (-r.y * p.x + r.x * p.y) * s);
}
// geometry.cu:69-75
inline f3 rotateBeam(f3 v, float theta, float phi, float coll) {
    float s, c;
    sincosf(-phi, &s, &c);
    f3 axis = mk(s, 0.f, c);
    f3 tmp = rot_origin(v, axis, -theta);
    return rot_origin(tmp, mk(0.f, 1.f, 0.f), phi + coll);
}
// geometry.cu:85-92
inline f3 invRotateBeam(f3 v, float theta, float phi, float coll) {
    f3 tmp = rot_origin(v, mk(0.f, 1.f, 0.f), -(phi + coll));
    float s, c;
    sincosf(-phi, &s, &c);
    f3 axis = mk(s, 0.f, c);
    return rot_origin(tmp, axis, theta);
}
// geometry.cu:110-115
inline f3 invRotateBeamlet(f3 v, float theta, float phi) {
    float s, c;
    sincosf(-phi, &s, &c);
    return rot_origin(v, mk(s, 0.f, c), -theta);
}
// geometry.cu:99-101
inline f3 rotateBeamlet(f3 v, float theta, float phi) { return invRotateBeamlet(v, -theta, phi); }
// geometry.cu:126-131
inline f3 rotateKernel(f3 v, float theta, float phi) {
    float s, c;
    sincosf(phi, &s, &c);
    f3 untilt = rot_origin(v, mk(c, 0.f, -s), -theta);
    return rot_origin(untilt, mk(0.f, 1.f, 0.f), -phi);
}
// geometry.cu:140-145
inline f3 invRotateKernel(f3 v, float theta, float phi) {
    f3 roll = rot_origin(v, mk(0.f, 1.f, 0.f), phi);
    float s, c;
    sincosf(phi, &s, &c);
    return rot_origin(roll, mk(c, 0.f, -s), theta);
}

// ---- explicit texture fetches ---------------------------------------------------------------
struct Vol {
    const float* d;
    int nx, ny, nz;
    inline float at(int i, int j, int k) const {
        if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return 0.f;
        return d[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)];
    }
};
// tex3D with cudaFilterModeLinear + cudaAddressModeBorder: texel i is centred at coordinate
// i + 0.5; taps outside the array read 0.  Weights in full FP32.
inline float tex3d(const Vol& v, float u, float w, float t) {
    float xb = u - 0.5f, yb = w - 0.5f, zb = t - 0.5f;
    if (!(xb > -1.f && yb > -1.f && zb > -1.f && xb < (float)v.nx && yb < (float)v.ny && zb < (float)v.nz))
        return 0.f;
    float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    int i = (int)fx, j = (int)fy, k = (int)fz;
    float a = xb - fx, b = yb - fy, c = zb - fz;
    float v000 = v.at(i, j, k), v100 = v.at(i + 1, j, k), v010 = v.at(i, j + 1, k), v110 = v.at(i + 1, j + 1, k);
    float v001 = v.at(i, j, k + 1), v101 = v.at(i + 1, j, k + 1), v011 = v.at(i, j + 1, k + 1), v111 = v.at(i + 1, j + 1, k + 1);
    return (1 - a) * (1 - b) * (1 - c) * v000 + a * (1 - b) * (1 - c) * v100 + (1 - a) * b * (1 - c) * v010 +
           a * b * (1 - c) * v110 + (1 - a) * (1 - b) * c * v001 + a * (1 - b) * c * v101 +
           (1 - a) * b * c * v011 + a * b * c * v111;
}
// tex2D with cudaFilterModeLinear + cudaAddressModeClamp over the kernel table
// (width nradii, height ntheta), as bound in nvbbRayConvolve_host.cu:182-184.
inline float tex2d_clamp(const float* K, int w, int h, float x, float y) {
    float xb = x - 0.5f, yb = y - 0.5f;
    float fx = floorf(xb), fy = floorf(yb);
    int i = (int)fx, j = (int)fy;
    float a = xb - fx, b = yb - fy;
    auto at = [&](int ii, int jj) {
        ii = std::min(std::max(ii, 0), w - 1);
        jj = std::min(std::max(jj, 0), h - 1);
        return K[ii + w * jj];
    };
    return (1 - a) * (1 - b) * at(i, j) + a * (1 - b) * at(i + 1, j) + (1 - a) * b * at(i, j + 1) + a * b * at(i + 1, j + 1);
}

struct Prob {
    const cccs_problem* p;
    Vol dens;
    f3 start, voxel, size_f;
    int nrays;
    float theta(int ray) const { return p->conv_theta_deg[ray] * kPI / 180.f; }   // brain_defs.h:67-73
    float phi(int ray) const { return p->conv_phi_deg[ray] * kPI / 180.f; }
#if ORC_MUTANT == 1
    int kern_index(int ray) const { return ray / (int)(p->ntheta / 2); }  // mutant: the phi index
#else
    int kern_index(int ray) const { return ray % (int)(p->ntheta / 2); }  // get_kernel_theta_index
#endif
    explicit Prob(const cccs_problem* pp) : p(pp) {
        dens = Vol{pp->density, (int)pp->size[0], (int)pp->size[1], (int)pp->size[2]};
        start = from_f(pp->start);
        voxel = from_f(pp->voxel);
        size_f = from_u3(pp->size);
        nrays = (int)(pp->nphi * (pp->ntheta / 2));
    }
};

// ---- A2: cudaRaytrace_device (DoseCalcAlgo/cudaSiddon.cu:20-115) ----------------------------
struct Seg { int32_t ijk[3]; float mid[3]; float len; };

template <class Cb>
float siddon_rpl(f3 this_voxel, f3 start, f3 end, f3 voxel, f3 src, const Vol& dens, Cb&& cb, uint64_t* steps) {
    f3 diff = this_voxel - src;
    f3 startdiff = start - src;
    f3 enddiff = end - src;
    float a_min, a_max;
    {
        f3 a_first, a_last;
        a_first.x = ((startdiff.x) - 0.5f * voxel.x) / (diff.x);
        a_first.y = ((startdiff.y) - 0.5f * voxel.y) / (diff.y);
        a_first.z = ((startdiff.z) - 0.5f * voxel.z) / (diff.z);
        a_last.x = ((enddiff.x) + 0.5f * voxel.x) / (diff.x);
        a_last.y = ((enddiff.y) + 0.5f * voxel.y) / (diff.y);
        a_last.z = ((enddiff.z) + 0.5f * voxel.z) / (diff.z);
        f3 amin = mk(fminf(a_first.x, a_last.x), fminf(a_first.y, a_last.y), fminf(a_first.z, a_last.z));
        f3 amax = mk(fmaxf(a_first.x, a_last.x), fmaxf(a_first.y, a_last.y), fmaxf(a_first.z, a_last.z));
        a_min = fmaxf(0, fmaxf(fmaxf(amin.x, amin.y), amin.z));
        a_max = fminf(1, fminf(fminf(amax.x, amax.y), amax.z));
    }
    float rpl = 0.0f;
    if (!(a_min >= a_max)) {
        float d12 = length(diff);
        float step_x = fabsf(voxel.x / diff.x);
        float step_y = fabsf(voxel.y / diff.y);
        float step_z = fabsf(voxel.z / diff.z);
        float alpha = a_min, alpha_x = a_min, alpha_y = a_min, alpha_z = a_min;
        float nextalpha = 0.f;
        const int max_iters = 5000;
        int iter = 0;
        while (alpha < a_max && ++iter < max_iters) {
            bool vx = (alpha_x >= 0.0f && alpha_x < 1.0f);
            bool vy = (alpha_y >= 0.0f && alpha_y < 1.0f);
            bool vz = (alpha_z >= 0.0f && alpha_z < 1.0f);
            if (!(vx || vy || vz)) break;
            if (vx && (!vy || alpha_x <= alpha_y) && (!vz || alpha_x <= alpha_z)) { nextalpha = alpha_x; alpha_x += step_x; }
            if (vy && (!vx || alpha_y <= alpha_x) && (!vz || alpha_y <= alpha_z)) { nextalpha = alpha_y; alpha_y += step_y; }
            if (vz && (!vx || alpha_z <= alpha_x) && (!vy || alpha_z <= alpha_y)) { nextalpha = alpha_z; alpha_z += step_z; }
            float intersection = fabsf(d12 * (nextalpha - alpha));
            if (intersection >= 0.001f) {
                float alpha_mid = (nextalpha + alpha) * 0.5f;
                float fx = (src.x + alpha_mid * diff.x - start.x) / voxel.x;
                float fy = (src.y + alpha_mid * diff.y - start.y) / voxel.y;
                float fz = (src.z + alpha_mid * diff.z - start.z) / voxel.z;
                rpl += intersection * tex3d(dens, fx, fy, fz);
                cb(fx, fy, fz, intersection);
            }
            alpha = nextalpha;
        }
        if (steps) *steps += (uint64_t)iter;
    }
    return rpl;
}

struct SpecT {
    int n;
    const float *flu, *en, *mu_en, *mu;
};

// ---- A2: cudaSiddon_device (cudaSiddon.cu:117-208) ------------------------------------------
float siddon_terma(f3 this_voxel, const float* fmap, f3 start, f3 end, f3 voxel, float bhc,
                   const cccs_beam& b, const SpecT& sp, const Vol& dens, int beamletnum, uint64_t* steps,
                   bool* active) {
    float terma = 0.f;
    f3 src = from_f(b.source);
    f3 diff = this_voxel - src;
    f3 normDir = normalize(from_f(b.direction));
    f3 projection = dot(diff, normDir) * normDir;
    f3 rejection = diff - projection;
    float trad = length(rejection) * fabsf(length(from_f(b.isocenter) - src) / length(projection));
    f3 normRejection = normalize(rejection);
    f3 proj_fmap = trad * normRejection;
    proj_fmap = rotateBeam(proj_fmap, b.azimuth, b.zenith, b.coll);
    f3 fc = proj_fmap;
    fc.x = (fc.x / b.beamlet_size[0]) + ((float)b.fmap_size[0] * 0.5f);
    fc.z = (fc.z / b.beamlet_size[1]) + ((float)b.fmap_size[1] * 0.5f);
    if (!(fc.x < 0 || fc.x >= (float)b.fmap_size[0] || fc.z < 0 || fc.z >= (float)b.fmap_size[1])) {
        int fidx = f2i_rd(fc.x) + (int)b.fmap_size[0] * f2i_rd(fc.z);
        if (!(fidx < 0 || fidx >= (int)(b.fmap_size[0] * b.fmap_size[1]) || (beamletnum >= 0 && fidx != beamletnum))) {
            float fluence = fmap[fidx];
            if (!(fluence <= 0.f)) {
                if (active) *active = true;
                float rpl = siddon_rpl(this_voxel, start, end, voxel, src, dens, [](float, float, float, float) {}, steps);
                float tsum = 0.f;
                for (int e = 0; e < sp.n; e++) {
                    float this_fluence = fluence * sp.flu[e];
                    float energy = sp.en[e];
                    float mu = sp.mu[e];
                    tsum += this_fluence * energy * mu * expf(-1.f * mu * rpl);
                }
                float d = length(this_voxel - src);
                float invsq = (b.sad * b.sad) / (d * d);
                terma = tsum * bhc * invsq;
            }
        }
    }
    return terma;
}

SpecT spec_of(const cccs_problem* p) { return SpecT{(int)p->nspec, p->spec_fluence, p->spec_energy, p->spec_mu_en, p->spec_mu}; }

inline f3 volume_end(const Prob& P) { return mk(P.start.x + P.size_f.x * P.voxel.x, P.start.y + P.size_f.y * P.voxel.y, P.start.z + P.size_f.z * P.voxel.z); }

// ---- A4: findREV (dosecalc-beam/nvbbRayConvolve_host.cu:75-155; beamlet :104-173) ----------
struct Rev { f3 min_c; uint32_t size[3]; };

inline void update_extents(f3& mn, f3& mx, const f3& pt) {
    if (pt.x < mn.x) mn.x = pt.x;
    if (pt.y < mn.y) mn.y = pt.y;
    if (pt.z < mn.z) mn.z = pt.z;
    if (pt.x > mx.x) mx.x = pt.x;
    if (pt.y > mx.y) mx.y = pt.y;
    if (pt.z > mx.z) mx.z = pt.z;
}

// returns false when the grid exceeds max_rev (ArraySizeError)
bool find_rev(Rev& rev, const cccs_problem* p, const uint32_t* max_rev, f3 lim_max, float theta, float phi) {
    f3 b_diff = lim_max;  // lim_min = 0
    f3 mn = mk(FLT_MAX, FLT_MAX, FLT_MAX);
    f3 mx = mk(FLT_MIN, FLT_MIN, FLT_MIN);
    for (int xo = 0; xo < 2; xo++)
        for (int yo = 0; yo < 2; yo++)
            for (int zo = 0; zo < 2; zo++) {
                f3 in = mk(b_diff.x * xo, b_diff.y * yo, b_diff.z * zo);
                update_extents(mn, mx, rotateKernel(in, theta, phi));
            }
    rev.size[0] = (uint32_t)ceilf(fabsf(mx.y - mn.y) / p->rev_longspacing);
    rev.size[1] = (uint32_t)ceilf(fabsf(mx.z - mn.z) / p->rev_latspacing);
    rev.size[2] = (uint32_t)ceilf(fabsf(mx.x - mn.x) / p->rev_latspacing);
    rev.min_c = mn;
    return !(rev.size[0] > max_rev[0] || rev.size[1] > max_rev[1] || rev.size[2] > max_rev[2]);
}

// beam path: max_rev_size is re-derived in initCudaConstandTex (nvbbRayConvolve_host.cu:248-258)
void beam_max_rev(const cccs_problem* p, uint32_t* out) {
    float min_vox = std::min(p->rev_latspacing, p->rev_longspacing);
    f3 bb = from_u3(p->bbox_size);
    unsigned int diag = (unsigned int)ceilf(length(bb) * p->voxel[0] / min_vox);
    unsigned int buffer = 100;
    unsigned int v = (unsigned int)(diag + 2 * (0.0f / min_vox + p->penumbra / min_vox) + buffer);
    out[0] = out[1] = out[2] = v;
}

// ---- A3: crop box (dosecalc-beam/nvbbRayConvolve_host.cu:429-499) ---------------------------
int crop_box(const cccs_problem* p, const cccs_beam& b, const float* terma, cccs_beam_geometry* g) {
    const int bx = (int)p->bbox_size[0], by = (int)p->bbox_size[1], bz = (int)p->bbox_size[2];
    f3 bmax = mk(FLT_MIN, FLT_MIN, FLT_MIN);
    f3 bmin = mk(FLT_MAX, FLT_MAX, FLT_MAX);
#pragma omp parallel
    {
        f3 lmax = bmax, lmin = bmin;
#pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < bz; z++)
            for (int y = 0; y < by; y++)
                for (int x = 0; x < bx; x++) {
                    size_t vid = (size_t)x + (size_t)bx * ((size_t)y + (size_t)by * z);
                    if (terma[vid] != 0) {
                        f3 bi = rotateBeam(mk((float)x, (float)y, (float)z), b.azimuth, b.zenith, b.coll);
                        lmin = mk(fminf(lmin.x, bi.x), fminf(lmin.y, bi.y), fminf(lmin.z, bi.z));
                        lmax = mk(fmaxf(lmax.x, bi.x), fmaxf(lmax.y, bi.y), fmaxf(lmax.z, bi.z));
                    }
                }
#pragma omp critical
        {
            bmin = mk(fminf(bmin.x, lmin.x), fminf(bmin.y, lmin.y), fminf(bmin.z, lmin.z));
            bmax = mk(fmaxf(bmax.x, lmax.x), fmaxf(bmax.y, lmax.y), fmaxf(bmax.z, lmax.z));
        }
    }
    bmin.x -= (p->penumbra / p->rev_latspacing);
    bmin.y -= (p->penumbra / p->rev_longspacing);
    bmin.z -= (p->penumbra / p->rev_latspacing);
    bmax.x += (p->penumbra / p->rev_latspacing);
    bmax.y += (p->penumbra / p->rev_longspacing);
    bmax.z += (p->penumbra / p->rev_latspacing);
    float sx = ceilf(fabsf(bmax.x - bmin.x)), sy = ceilf(fabsf(bmax.y - bmin.y)), sz = ceilf(fabsf(bmax.z - bmin.z));
    // the reference converts to uint3 and rejects sizes <= 0 or >= 1000 (:494-498)
    if (!(sx >= 1.f && sx < 1000.f && sy >= 1.f && sy < 1000.f && sz >= 1.f && sz < 1000.f)) return CCCS_EGEOM;
    g->bev_size[0] = (uint32_t)sx; g->bev_size[1] = (uint32_t)sy; g->bev_size[2] = (uint32_t)sz;
    g->bev_min[0] = bmin.x; g->bev_min[1] = bmin.y; g->bev_min[2] = bmin.z;
    g->bev_max[0] = bmax.x; g->bev_max[1] = bmax.y; g->bev_max[2] = bmax.z;
    f3 tmin = invRotateBeam(bmin, b.azimuth, b.zenith, b.coll);
    f3 st = from_f(p->start), vx = from_f(p->voxel), bs = from_u3(p->bbox_start);
    g->lim_min[0] = st.x + (tmin.x + bs.x) * vx.x;
    g->lim_min[1] = st.y + (tmin.y + bs.y) * vx.y;
    g->lim_min[2] = st.z + (tmin.z + bs.z) * vx.z;
    return CCCS_OK;
}

int beam_rev_grids(const cccs_problem* p, cccs_beam_geometry* g) {
    Prob P(p);
    if (P.nrays > 48) return CCCS_EINVAL;
    uint32_t mr[3];
    beam_max_rev(p, mr);
    f3 rvs = mk(p->rev_latspacing, p->rev_longspacing, p->rev_latspacing);
    f3 lim_max = from_u3(g->bev_size) * rvs;
    g->nrays = P.nrays;
    int status = CCCS_OK;
    for (int r = 0; r < P.nrays; r++) {
        Rev rev;
        if (!find_rev(rev, p, mr, lim_max, P.theta(r), P.phi(r))) status = CCCS_EGEOM;
        if (rev.size[0] > 1024) status = CCCS_EGEOM;  // :567
        g->rev_min[r][0] = rev.min_c.x; g->rev_min[r][1] = rev.min_c.y; g->rev_min[r][2] = rev.min_c.z;
        for (int k = 0; k < 3; k++) g->rev_size[r][k] = rev.size[k];
    }
    return status;
}

// ---- A6: RowConvolve line body (dosecalc-beam/nvbbRayConvolve_device.cu:113-217) -----------
// rows: kernel table (ntheta x nradii), R = KERN_RADII (radial boundaries, host.cu:224)
float row_convolve_one(const float* dens, const float* terma, int n, int tx, const float* K, int ntheta,
                       int nradii, const float* R, float P, float delr, int nphi, uint64_t* it_valid,
                       uint64_t* it_all) {
    if (dens[tx] < 0.0f) return 0.f;
    float sum = 0.f, r_eff = 0.f, delr_eff = 0.f, cumval = 0.f, last_cumval = 0.f;
    int rad_start = 0, k1 = tx;
    float last_rad = 0.f, next_rad = R[0];
    do {
        k1--;
        if (k1 < 0) break;
        delr_eff = delr * dens[k1];
        r_eff += 0.5f * delr_eff;
        if (r_eff > R[nradii - 1]) break;
        while (next_rad < r_eff) {
            if (rad_start == nradii - 1) break;
            last_rad = next_rad;
            next_rad = R[++rad_start];
        }
        float interp = (r_eff - last_rad) / (next_rad - last_rad);
        if (rad_start == 0) cumval = interp * tex2d_clamp(K, nradii, ntheta, 0.5f, P + 0.5f);
        else cumval = tex2d_clamp(K, nradii, ntheta, (float)rad_start - 0.5f + interp, P + 0.5f);
        float t = terma[k1];
        sum += (cumval - last_cumval) * t;
        last_cumval = cumval;
        r_eff += 0.5f * delr_eff;
        if (it_all) (*it_all)++;
        if (it_valid && dens[k1] >= 0.f) (*it_valid)++;
    } while (rad_start < nradii - 1);
    r_eff = delr_eff = 0.f;
    cumval = last_cumval = 0.f;
    rad_start = 0;
    k1 = tx;
    last_rad = 0.f;
    next_rad = R[0];
    do {
        k1++;
        if (k1 >= n) break;
        delr_eff = delr * dens[k1];
        r_eff += 0.5f * delr_eff;
        if (r_eff > R[nradii - 1]) break;
        while (next_rad < r_eff) {
            if (rad_start == nradii - 1) break;
            last_rad = next_rad;
            next_rad = R[++rad_start];
        }
        float interp = (r_eff - last_rad) / (next_rad - last_rad);
        if (rad_start == 0) cumval = interp * tex2d_clamp(K, nradii, ntheta, 0.5f, (float)ntheta - 0.5f - P);
        else cumval = tex2d_clamp(K, nradii, ntheta, (float)rad_start - 0.5f + interp, (float)ntheta - 0.5f - P);
        float t = terma[k1];
        sum += (cumval - last_cumval) * t;
        last_cumval = cumval;
        r_eff += 0.5f * delr_eff;
        if (it_all) (*it_all)++;
        if (it_valid && dens[k1] >= 0.f) (*it_valid)++;
    } while (rad_start < nradii - 1);
    return sum / (float)nphi;
}

// ---- A12: PackRowConvolve line body (dosecalc-beamlet/nvbbRayConvolve_device.cu:284-417) ---
// R = KERN_RADII (mean radii, dosecalc-beamlet/nvbbRayConvolve_host.cu:351-357)
float pack_row_convolve_one(const float* dens, const float* terma, int n, int tx, const float* K, int ntheta,
                            int nradii, const float* R, float Pf, float delr, int nphi, uint64_t* iters) {
    if (dens[tx] < 0.0f) return 0.f;
    float P = Pf;
    float recip_P = (float)ntheta - Pf - 1;
    float sum = 0.f;
    int kk = tx;
    float r_eff = 0.f;
    int idx_rad = 0;
    float last_cumval = 0.f;
    do {
        if (kk < 0) break;
        if (dens[kk] < 0.f) break;
        float delr_eff = delr * dens[kk];
        r_eff += 0.5f * delr_eff;
        if (r_eff > R[nradii - 1]) break;
        while (R[idx_rad] < r_eff) idx_rad++;
        float cumval;
        if (idx_rad <= 0) {
            float interp = r_eff * (1.0f / R[0]);
            cumval = tex2d_clamp(K, nradii, ntheta, 0.5f, P + 0.5f) * interp;
        } else {
            float interp = (r_eff - R[idx_rad - 1]) * (1.0f / (R[idx_rad] - R[idx_rad - 1]));
            cumval = tex2d_clamp(K, nradii, ntheta, (float)idx_rad - 0.5f + interp, P + 0.5f);
        }
        sum += (cumval - last_cumval) * terma[kk];
        last_cumval = cumval;
        if (kk != tx) r_eff += 0.5f * delr_eff;
        --kk;
        if (iters) (*iters)++;
    } while (idx_rad < nradii - 1);
    kk = tx;
    r_eff = 0.f;
    idx_rad = 0;
    last_cumval = 0.f;
    do {
        if (kk >= n) break;
        if (dens[kk] < 0.f) break;
        float delr_eff = delr * dens[kk];
        r_eff += 0.5f * delr_eff;
        if (r_eff > R[nradii - 1]) break;
        while (R[idx_rad] < r_eff) idx_rad++;
        float cumval;
        if (idx_rad <= 0) {
            float interp = r_eff * (1.0f / R[0]);
            cumval = tex2d_clamp(K, nradii, ntheta, 0.5f, recip_P + 0.5f) * interp;
        } else {
            float interp = (r_eff - R[idx_rad - 1]) * (1.0f / (R[idx_rad] - R[idx_rad - 1]));
            cumval = tex2d_clamp(K, nradii, ntheta, (float)idx_rad - 0.5f + interp, recip_P + 0.5f);
        }
        sum += (cumval - last_cumval) * terma[kk];
        last_cumval = cumval;
        if (kk != tx) r_eff += 0.5f * delr_eff;
        ++kk;
        if (iters) (*iters)++;
    } while (idx_rad < nradii - 1);
    return sum / (float)nphi;
}

// full-beam TERMA over the calc_bbox (cudaSiddon kernel, cudaSiddon.cu:210-308)
void terma_bbox(const cccs_problem* p, const cccs_beam& b, const float* fmap, int beamletnum, float* out,
                orc_stats* st) {
    Prob P(p);
    SpecT sp = spec_of(p);
    f3 end = volume_end(P);
    const int bx = (int)p->bbox_size[0], by = (int)p->bbox_size[1], bz = (int)p->bbox_size[2];
    const int ss = std::max(1, (int)p->ss_factor);
    const int nsub = ss * ss * ss;
    uint64_t steps = 0, active = 0;
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : steps, active)
    for (int Z = 0; Z < bz; Z++)
        for (int Y = 0; Y < by; Y++)
            for (int X = 0; X < bx; X++) {
                float acc = 0.f;
                bool act = false;
                if (nsub > 1) {
                    // supersampled TERMA, averaged in thread order (cudaSiddon.cu:249-260, 300-306)
                    f3 subsz = P.voxel / (float)ss;
                    float sub[512];
                    for (int s = 0; s < nsub; s++) {
                        int sX = s % ss, sY = (s / ss) % ss, sZ = s / (ss * ss);
                        f3 v = mk(P.start.x + ((float)p->bbox_start[0] + X - 0.5f) * P.voxel.x + (sX + 0.5f) * subsz.x,
                                  P.start.y + ((float)p->bbox_start[1] + Y - 0.5f) * P.voxel.y + (sY + 0.5f) * subsz.y,
                                  P.start.z + ((float)p->bbox_start[2] + Z - 0.5f) * P.voxel.z + (sZ + 0.5f) * subsz.z);
                        sub[s] = siddon_terma(v, fmap, P.start, end, P.voxel, p->beamhard_correct, b, sp, P.dens,
                                              beamletnum, &steps, &act);
                    }
                    for (int s = 0; s < nsub; s++) acc += sub[s];
                    acc /= (float)nsub;
                } else {
                    f3 v = mk(P.start.x + ((float)p->bbox_start[0] + X) * P.voxel.x,
                              P.start.y + ((float)p->bbox_start[1] + Y) * P.voxel.y,
                              P.start.z + ((float)p->bbox_start[2] + Z) * P.voxel.z);
                    acc = siddon_terma(v, fmap, P.start, end, P.voxel, p->beamhard_correct, b, sp, P.dens, beamletnum,
                                       &steps, &act);
                }
                if (act) active++;
                out[(size_t)X + (size_t)bx * ((size_t)Y + (size_t)by * Z)] = acc;
            }
    if (st) {
        st->terma_voxels += (uint64_t)bx * by * bz;
        st->terma_active += active;
        st->siddon_steps += steps;
    }
}

// A5 + A6 for one cone: REV-grid sized dens/terma/dose (x = line fastest)
void beam_cone(const cccs_problem* p, const cccs_beam& b, const cccs_beam_geometry& g, int ray,
               const float* terma, float* rdens, float* rterma, float* rdose, orc_stats* st) {
    Prob P(p);
    const uint32_t* rs = g.rev_size[ray];
    const int nx = (int)rs[0], ny = (int)rs[1], nz = (int)rs[2];
    f3 revStart = from_f(g.rev_min[ray]);
    f3 rvs = mk(p->rev_latspacing, p->rev_longspacing, p->rev_latspacing);
    f3 rrv = rcp3(rvs);
    f3 bevStart = from_f(g.lim_min);
    f3 bsz = from_u3(g.bev_size);
    float kt = P.theta(ray), kp = P.phi(ray);
    f3 bb0 = from_u3(p->bbox_start);
    f3 bb1 = bb0 + from_u3(p->bbox_size);
    Vol tv{terma, (int)p->bbox_size[0], (int)p->bbox_size[1], (int)p->bbox_size[2]};
    f3 off = (bevStart - P.start);
    f3 rvox = rcp3(P.voxel);
    uint64_t valid = 0;
    // cudaRaytrace (dosecalc-beam/nvbbRayConvolve_device.cu:11-80)
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : valid)
    for (int tZ = 0; tZ < nz; tZ++)
        for (int tY = 0; tY < ny; tY++)
            for (int tX = 0; tX < nx; tX++) {
                size_t o = (size_t)tX + (size_t)nx * ((size_t)tY + (size_t)ny * tZ);
                rterma[o] = 0.f;
                f3 bev = invRotateKernel(fma3(revStart, rrv, mk((float)tZ, (float)tX, (float)tY)), kt, kp);
                if (bev.x < 0.f || bev.y < 0.f || bev.z < 0.f || bev.x >= bsz.x || bev.y >= bsz.y || bev.z >= bsz.z) {
                    rdens[o] = -3.f;
                    continue;
                }
                f3 gi = fma3(off, rvox, invRotateBeam(bev, b.azimuth, b.zenith, b.coll));
                if (gi.x < bb0.x || gi.x >= bb1.x || gi.y < bb0.y || gi.y >= bb1.y || gi.z < bb0.z || gi.z >= bb1.z) {
                    rdens[o] = -2.f;
                    continue;
                }
                rdens[o] = tex3d(P.dens, gi.x + 0.5f, gi.y + 0.5f, gi.z + 0.5f);
                f3 ti = gi - bb0;
                rterma[o] = tex3d(tv, ti.x + 0.5f, ti.y + 0.5f, ti.z + 0.5f);
                valid++;
            }
    // RowConvolve (:87-218)
    const float* R = p->radii;
    int Pidx = P.kern_index(ray);
    uint64_t itv = 0, ita = 0;
#pragma omp parallel for collapse(2) schedule(dynamic, 8) reduction(+ : itv, ita)
    for (int tZ = 0; tZ < nz; tZ++)
        for (int tY = 0; tY < ny; tY++) {
            size_t base = (size_t)nx * ((size_t)tY + (size_t)ny * tZ);
            for (int tX = 0; tX < nx; tX++)
                rdose[base + tX] = row_convolve_one(rdens + base, rterma + base, nx, tX, p->kernel, (int)p->ntheta,
                                                    (int)p->nradii, R, (float)Pidx, p->rev_longspacing,
                                                    (int)p->nphi, &itv, &ita);
        }
    if (st) {
        st->rev_samples += (uint64_t)nx * ny * nz;
        st->rev_valid += valid;
        st->cone_iters += itv;
        st->cone_iters_all += ita;
    }
}

// A7: REV2XYZdose (dosecalc-beam/nvbbRayConvolve_device.cu:222-261), dose += per voxel
void beam_rev_to_pvcs(const cccs_problem* p, const cccs_beam& b, const cccs_beam_geometry& g, int ray,
                      const float* rdose, float* dose, orc_stats* st) {
    Prob P(p);
    const uint32_t* rs = g.rev_size[ray];
    Vol rv{rdose, (int)rs[0], (int)rs[1], (int)rs[2]};
    f3 revStart = from_f(g.rev_min[ray]);
    f3 rvs = mk(p->rev_latspacing, p->rev_longspacing, p->rev_latspacing);
    f3 rrv = rcp3(rvs);
    f3 bevStart = from_f(g.lim_min);
    float kt = P.theta(ray), kp = P.phi(ray);
    f3 rvox = rcp3(P.voxel);
    const int x0 = (int)p->bbox_start[0], y0 = (int)p->bbox_start[1], z0 = (int)p->bbox_start[2];
    const int x1 = x0 + (int)p->bbox_size[0], y1 = y0 + (int)p->bbox_size[1], z1 = z0 + (int)p->bbox_size[2];
    const size_t NX = p->size[0], NY = p->size[1];
    uint64_t hits = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : hits)
    for (int tZ = z0; tZ < z1; tZ++)
        for (int tY = y0; tY < y1; tY++)
            for (int tX = x0; tX < x1; tX++) {
                f3 gc = fma3(mk((float)tX, (float)tY, (float)tZ), P.voxel, P.start);
                f3 bi = rotateBeam((gc - bevStart) * rvox, b.azimuth, b.zenith, b.coll);
                f3 ri = fma3(-revStart, rrv, rotateKernel(bi, kt, kp));
                if (ri.y >= (float)rs[0] || ri.z >= (float)rs[1] || ri.x >= (float)rs[2] || ri.x < 0.f || ri.y < 0.f ||
                    ri.z < 0.f)
                    continue;
                hits++;
                dose[(size_t)tX + NX * ((size_t)tY + NY * tZ)] += tex3d(rv, ri.y + 0.5f, ri.z + 0.5f, ri.x + 0.5f);
            }
    if (st) st->pvcs_hits += hits;
}

int beam_geometry_full(const cccs_problem* p, const cccs_beam& b, const float* terma, cccs_beam_geometry* g) {
    std::memset(g, 0, sizeof(*g));
    int s = crop_box(p, b, terma, g);
    if (s != CCCS_OK) return s;
    return beam_rev_grids(p, g);
}

bool problem_ok(const cccs_problem* p) {
    if (!p || !p->density || !p->kernel || !p->radii || !p->conv_theta_deg || !p->conv_phi_deg) return false;
    if (p->ntheta < 2 || (p->ntheta % 2) || p->nphi < 1 || p->nradii < 2 || p->nspec < 1 || p->nspec > 30) return false;
    for (int k = 0; k < 3; k++)
        if (p->size[k] == 0 || p->bbox_size[k] == 0 || p->bbox_start[k] + p->bbox_size[k] > p->size[k]) return false;
    return true;
}

// ---- A10: calcBeamletAnchors (dosecalc-beamlet/nvbbRayConvolve_host.cu:180-280) -------------
bool beamlet_anchors(f3& start, f3& end, f2& angles, f3& biso, f3 src, f3 iso, unsigned idx, f2 bsize,
                     const uint32_t* fdims, f3 voxel, f3 dstart, const uint32_t* bb_start, const uint32_t* bb_size,
                     float gantry, float couch, float coll) {
    unsigned bx = idx % fdims[0], by = idx / fdims[0];
    f3 bdiff = mk(bsize.x * (-0.5f * (float)fdims[0] + (float)bx + 0.5f), 0,
                  bsize.y * (-0.5f * (float)fdims[1] + (float)by + 0.5f));
    biso = iso + invRotateBeam(bdiff, gantry, couch, coll);
    angles = f2{acosf(length(iso - src) / length(biso - src)), atan2f(bdiff.z, bdiff.x)};
    f3 bbox_begin = dstart + voxel * from_u3(bb_start);
    f3 bbox_end = bbox_begin + voxel * from_u3(bb_size);
    double buffer = 0.05;
    f3 inter[6];
    start = mk(FLT_MAX, FLT_MAX, FLT_MAX);
    end = mk(FLT_MIN, FLT_MIN, FLT_MIN);
    double min_dist = std::numeric_limits<double>::max();
    double max_dist = std::numeric_limits<double>::min();
    for (int ii = 0; ii < 3; ii++) {
        f3 d = biso - src;
        double denom = comp(d, ii);
        double a1 = (comp(bbox_begin, ii) - comp(src, ii)) / (denom + 1e-12);
        double a2 = (comp(bbox_end, ii) - comp(src, ii)) / (denom + 1e-12);
        inter[2 * ii] = src + (float)a1 * d;
        inter[2 * ii + 1] = src + (float)a2 * d;
    }
    for (int ii = 0; ii < 3; ii++)
        for (int jj = 0; jj < 2; jj++) {
            const f3& it = inter[2 * ii + jj];
            int i1 = (ii + 1) % 3, i2 = (ii + 2) % 3;
            if (comp(it, i1) + buffer >= comp(bbox_begin, i1) && comp(it, i1) - buffer <= comp(bbox_end, i1) &&
                comp(it, i2) + buffer >= comp(bbox_begin, i2) && comp(it, i2) - buffer <= comp(bbox_end, i2)) {
                f3 in = it;
                if (length(in - src) < min_dist) { start = in; min_dist = length(start - src); }
                if (length(in - src) > max_dist) { end = in; max_dist = length(end - src); }
            }
        }
    return (min_dist < std::numeric_limits<float>::max() && max_dist > std::numeric_limits<float>::min());
}

struct PillarGrid {
    int n = 0;
    int npx = 0, npy = 0;
    int gd[3] = {0, 0, 0};
    int pd[3] = {0, 0, 0};
    f3 psize{0, 0, 0};
    int wall = 2, buffer = 2;
    std::vector<int> idx;
    std::vector<f3> pstart;
    std::vector<f2> angles;
};

// Test hooks of the beamlet path (the oracle's own instrumentation, not reference behaviour).
struct BeamletHooks {
    // stage dump: REV density, TERMA and cone dose of batch 0 for this cone, then stop
    int stage_ray = -1;
    float *rd = nullptr, *rt = nullptr, *rdo = nullptr;
    uint64_t cap = 0;
    uint32_t* rev_size = nullptr;
    // linear field: every packed node gets linear[0] + (linear[1], linear[2], linear[3]) . X(node), X the
    // PVCS point the raytrace assigns to the node, instead of the cone dose; then the unpack runs
    const float* linear = nullptr;
};

// PVCS point of pillar-local REV coordinate inp (the density point of cudaBeamletRaytrace,
// dosecalc-beamlet/nvbbRayConvolve_device.cu:213-219)
inline f3 pillar_point(f3 inp, f3 rvs, f2 a, const cccs_beam& b, f3 pst) {
    return invRotateBeam(invRotateBeamlet(inp * rvs, a.x, a.y), b.azimuth, b.zenith, b.coll) + pst;
}

// dosecalc-beamlet radconvolveTexture (nvbbRayConvolve_host.cu:555-1349) for one beam.
int beamlet_run(const cccs_problem* p, const cccs_beam& b, const float* fmap, float* out, int max_beamlets,
                int* nactive, orc_stats* st, const BeamletHooks& hk) {
    Prob P(p);
    SpecT sp = spec_of(p);
    f3 end_vol = volume_end(P);
    f3 rvs = mk(p->rev_latspacing, p->rev_longspacing, p->rev_latspacing);
    f3 rrv = rcp3(rvs);
    f3 src = from_f(b.source), iso = from_f(b.isocenter);
    // KERN_RADII = mean radii (dosecalc-beamlet/nvbbRayConvolve_host.cu:351-357)
    std::vector<float> R(p->nradii);
#if ORC_MUTANT == 2
    for (uint32_t r = 0; r < p->nradii; r++) R[r] = p->radii[r];  // mutant: the beam path's raw radii
#else
    R[0] = (float)(0.5 * p->radii[0]);
    for (uint32_t r = 1; r < p->nradii; r++) R[r] = (float)(0.5 * (p->radii[r] + p->radii[r - 1]));
#endif
    // active beamlets (:575-581)
    PillarGrid H;
    for (uint32_t i = 0; i < b.fmap_size[0] * b.fmap_size[1]; i++)
        if (fmap[i] > 0) H.idx.push_back((int)i);
    H.n = (int)H.idx.size();
    if (nactive) *nactive = H.n;
    if (H.n == 0) return hk.stage_ray >= 0 ? CCCS_EINVAL : CCCS_OK;
    const bool stage = hk.stage_ray >= 0;
    if (!stage && H.n > max_beamlets) return CCCS_EINVAL;
    const size_t nbb = (size_t)p->bbox_size[0] * p->bbox_size[1] * p->bbox_size[2];
    if (!stage) std::memset(out, 0, sizeof(float) * nbb * H.n);
    // anchors, lengths, divergence sizes (:609-636)
    std::vector<f3> bstart(H.n), bend(H.n);
    std::vector<float> blen(H.n);
    H.angles.resize(H.n);
    float max_len = 0.f;
    f2 max_bs{0.f, 0.f};
    f2 bsz{b.beamlet_size[0], b.beamlet_size[1]};
    for (int ii = 0; ii < H.n; ii++) {
        f3 biso;
        if (!beamlet_anchors(bstart[ii], bend[ii], H.angles[ii], biso, src, iso, (unsigned)H.idx[ii], bsz, b.fmap_size,
                             P.voxel, P.start, p->bbox_start, p->bbox_size, b.azimuth, b.zenith, b.coll))
            return CCCS_EGEOM;
        blen[ii] = length(bend[ii] - bstart[ii]);
        if (blen[ii] > max_len) max_len = blen[ii];
        float ratio_num = length(bend[ii] - src), ratio_den = length(biso - src);
        f2 dv{bsz.x * ratio_num / ratio_den, bsz.y * ratio_num / ratio_den};
        max_bs.x = std::max(max_bs.x, dv.x);
        max_bs.y = std::max(max_bs.y, dv.y);
    }
    // pillar size (:638-649)
    float psize_long = max_len + 2.f * p->kernel_extent + H.wall * rvs.y;
    f2 psize_trans{max_bs.x + 2.f * p->kernel_extent + H.wall * rvs.x, max_bs.y + 2.f * p->kernel_extent + H.wall * rvs.x};
    f3 expand = rvs - mk(fmodf(psize_trans.x, rvs.x), fmodf(psize_long, rvs.y), fmodf(psize_trans.y, rvs.z));
    H.psize = mk(psize_trans.x, psize_long, psize_trans.y) + expand + 2.f * float(H.buffer) * rvs;
    f3 pdf = H.psize / rvs;
    H.pd[0] = (int)pdf.x; H.pd[1] = (int)pdf.y; H.pd[2] = (int)pdf.z;
    // pillar start coords (:652-661)
    H.pstart.resize(H.n);
    for (int ii = 0; ii < H.n; ii++) {
        f3 go = mk(-0.5f * (H.psize.x + H.wall * rvs.x), -0.5f * (H.psize.y - blen[ii] + H.wall * rvs.y),
                   -0.5f * (H.psize.z + H.wall * rvs.z));
        go = invRotateBeam(invRotateBeamlet(go, H.angles[ii].x, H.angles[ii].y), b.azimuth, b.zenith, b.coll);
        H.pstart[ii] = bstart[ii] + go;
    }
    const uint32_t* max_rev = p->max_rev_size;
    uint32_t mr_default[3] = {500, 500, 500};  // argparser default (dosecalc-preprocess/argparser.h:29)
    if (max_rev[0] == 0) max_rev = mr_default;
    // batching with retries on ArraySizeError (:692-834, allocate_batch_sizes :69-80)
    int nb = 1;
    std::vector<int> bsizes;
    for (;;) {
        if (nb > 100) return CCCS_ENOMEM;
        bsizes.assign(nb, H.n / nb);
        int rem = H.n - (H.n / nb) * nb, mk_ = 0;
        while (rem > 0) { bsizes[(mk_++) % nb]++; rem--; }
        bool ok = true;
        for (int bi = 0; bi < nb && ok; bi++) {
            int n = bsizes[bi];
            int npx = (int)ceilf(sqrtf((float)n));
            int npy = (int)ceilf((float)n / npx);
            f3 gsize = mk((float)(npx * H.pd[0] + H.wall), (float)(H.pd[1] + H.wall), (float)(npy * H.pd[2] + H.wall)) * rvs;
            for (int r = 0; r < P.nrays && ok; r++) {
                Rev rv;
                if (!find_rev(rv, p, max_rev, gsize, P.theta(r), P.phi(r))) ok = false;
                else if (rv.size[0] > 1024) return CCCS_EGEOM;
            }
        }
        if (ok) break;
        nb++;
    }
    int memidx = 0;
    for (int bi = 0; bi < nb; bi++) {
        const int n = bsizes[bi];
        const int* bidx = H.idx.data() + memidx;
        const f3* pst = H.pstart.data() + memidx;
        const f2* ang = H.angles.data() + memidx;
        int npx = (int)ceilf(sqrtf((float)n));
        int npy = (int)ceilf((float)n / npx);
        int gd[3] = {npx * H.pd[0] + H.wall, H.pd[1] + H.wall, npy * H.pd[2] + H.wall};
        f3 gsize = mk((float)gd[0], (float)gd[1], (float)gd[2]) * rvs;
        f3 gdf = mk((float)gd[0], (float)gd[1], (float)gd[2]);
        f3 pdf3 = mk((float)H.pd[0], (float)H.pd[1], (float)H.pd[2]);
        // packed BEV dose, layout [pX][pZ][pY] (PackedREVtoBEVdose, device.cu:422)
        std::vector<float> pdose((size_t)gd[0] * gd[1] * gd[2], 0.f);
        for (int ray = 0; ray < P.nrays && !hk.linear; ray++) {
            if (stage && ray != hk.stage_ray) continue;
            Rev rev;
            find_rev(rev, p, max_rev, gsize, P.theta(ray), P.phi(ray));
            const int nx = (int)rev.size[0], ny = (int)rev.size[1], nz = (int)rev.size[2];
            float kt = P.theta(ray), kp = P.phi(ray);
            size_t nrev = (size_t)nx * ny * nz;
            if (stage) {
                for (int k = 0; k < 3; k++) hk.rev_size[k] = rev.size[k];
                if (nrev > hk.cap || !hk.rd || !hk.rt || !hk.rdo) return CCCS_EINVAL;
            }
            std::vector<float> rd(nrev), rt(nrev), rdo(nrev);
            uint64_t valid = 0, steps = 0;
            // cudaBeamletRaytrace (device.cu:19-242), ss_factor forced to 1 (host.cu:747)
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : valid, steps)
            for (int tZ = 0; tZ < nz; tZ++)
                for (int tY = 0; tY < ny; tY++)
                    for (int tX = 0; tX < nx; tX++) {
                        size_t o = (size_t)tX + (size_t)nx * ((size_t)tY + (size_t)ny * tZ);
                        float vd = -1e-12f, vt = 0.f;
                        f3 pb = invRotateKernel(fma3(rev.min_c, rrv, mk((float)tZ, (float)tX, (float)tY)), kt, kp) + 0.5f;
                        bool act = !(pb.x < 0.f || pb.y < 0.f || pb.z < 0.f || pb.x >= gdf.x || pb.y >= gdf.y || pb.z >= gdf.z);
                        int pidx = 0;
                        f3 inp{0, 0, 0};
                        if (act) {
                            int pix = f2i_rd(pb.x / pdf3.x), piy = f2i_rd(pb.z / pdf3.z);
                            pidx = pix + piy * npx;
                            if (pidx >= n) act = false;
                        }
                        if (act) {
                            inp = mk(fmodf(pb.x, pdf3.x), fmodf(pb.y, pdf3.y), fmodf(pb.z, pdf3.z));
                            if (inp.x < H.wall || inp.y < H.wall || inp.z < H.wall) act = false;
                        }
                        if (act) {
                            inp = inp - 0.5f;
                            f2 a = ang[pidx];
                            f3 gcs = invRotateBeam(invRotateBeamlet(inp, a.x, a.y), b.azimuth, b.zenith, b.coll);
                            gcs = fma3(gcs, rvs, pst[pidx]);
                            vt = siddon_terma(gcs, fmap, P.start, end_vol, P.voxel, p->beamhard_correct, b, sp, P.dens,
                                              bidx[pidx], &steps, nullptr);
                            f3 gi = (pillar_point(inp, rvs, a, b, pst[pidx]) - P.start) / P.voxel;
                            vd = tex3d(P.dens, gi.x + 0.5f, gi.y + 0.5f, gi.z + 0.5f);
                            valid++;
                        }
                        rd[o] = vd;
                        rt[o] = vt;
                    }
            // PackRowConvolve (device.cu:253-418)
            uint64_t iters = 0;
            int Pk = P.kern_index(ray);
#pragma omp parallel for collapse(2) schedule(dynamic, 8) reduction(+ : iters)
            for (int tZ = 0; tZ < nz; tZ++)
                for (int tY = 0; tY < ny; tY++) {
                    size_t base = (size_t)nx * ((size_t)tY + (size_t)ny * tZ);
                    for (int tX = 0; tX < nx; tX++)
                        rdo[base + tX] = pack_row_convolve_one(rd.data() + base, rt.data() + base, nx, tX, p->kernel,
                                                               (int)p->ntheta, (int)p->nradii, R.data(), (float)Pk,
                                                               p->rev_longspacing, (int)p->nphi, &iters);
                }
            if (stage) {
                std::memcpy(hk.rd, rd.data(), nrev * sizeof(float));
                std::memcpy(hk.rt, rt.data(), nrev * sizeof(float));
                std::memcpy(hk.rdo, rdo.data(), nrev * sizeof(float));
                return CCCS_OK;
            }
            // PackedREVtoBEVdose (device.cu:427-452)
            Vol rv{rdo.data(), nx, ny, nz};
#pragma omp parallel for collapse(2) schedule(static)
            for (int pX = 0; pX < gd[0]; pX++)
                for (int pZ = 0; pZ < gd[2]; pZ++)
                    for (int pY = 0; pY < gd[1]; pY++) {
                        f3 ri = fma3(-rev.min_c, rrv, rotateKernel(mk((float)pX, (float)pY, (float)pZ), kt, kp));
                        float dc = tex3d(rv, ri.y + 0.5f, ri.z + 0.5f, ri.x + 0.5f);
                        if (dc <= 0.f) continue;
                        pdose[(size_t)pY + (size_t)gd[1] * ((size_t)pZ + (size_t)gd[2] * pX)] += dc;
                    }
            if (st) {
                st->rev_samples += nrev;
                st->rev_valid += valid;
                st->siddon_steps += steps;
                st->cone_iters += iters;
                st->cone_iters_all += iters;
            }
        }
        if (stage) return CCCS_EINVAL;  // unreachable: batch 0 holds every ray
        if (hk.linear) {
            // node (pX, pY, pZ) of pillar (pix, piy) is pillar-local REV coordinate (pX - pix*pd, pY, pZ - piy*pd)
            const float* c = hk.linear;
            for (int pX = 0; pX < gd[0]; pX++)
                for (int pZ = 0; pZ < gd[2]; pZ++) {
                    int pix = pX / H.pd[0], piy = pZ / H.pd[2], pi = pix + piy * npx;
                    if (pix >= npx || pi >= n) continue;
                    for (int pY = 0; pY < gd[1]; pY++) {
                        f3 x = pillar_point(mk((float)(pX - pix * H.pd[0]), (float)pY, (float)(pZ - piy * H.pd[2])), rvs,
                                            ang[pi], b, pst[pi]);
                        pdose[(size_t)pY + (size_t)gd[1] * ((size_t)pZ + (size_t)gd[2] * pX)] =
                            c[0] + c[1] * x.x + c[2] * x.y + c[3] * x.z;
                    }
                }
        }
        // UnpackBEVDosePillar (device.cu:459-521), one dense calc_bbox column per beamlet
        Vol pv{pdose.data(), gd[1], gd[2], gd[0]};
        const int bx = (int)p->bbox_size[0], by = (int)p->bbox_size[1], bz = (int)p->bbox_size[2];
        f3 fbs = from_u3(p->bbox_start);
        const int lo = H.wall + H.buffer;
        for (int pi = 0; pi < n; pi++) {
            float* col = out + (size_t)(memidx + pi) * nbb;
            int pix = pi % npx, piy = pi / npx;
            f2 a = ang[pi];
            f3 ps = pst[pi];
#pragma omp parallel for collapse(2) schedule(static)
            for (int tZ = 0; tZ < bz; tZ++)
                for (int tY = 0; tY < by; tY++)
                    for (int tX = 0; tX < bx; tX++) {
                        f3 gc = (mk((float)tX, (float)tY, (float)tZ) + fbs) * P.voxel + P.start;
#if ORC_MUTANT == 3
                        // mutant: the half-texel shift dropped along the beamlet axis
                        f3 pidx = rotateBeamlet(rotateBeam(gc - ps, b.azimuth, b.zenith, b.coll), a.x, a.y) / rvs +
                                  mk(0.5f, 0.f, 0.5f);
#else
                        f3 pidx = rotateBeamlet(rotateBeam(gc - ps, b.azimuth, b.zenith, b.coll), a.x, a.y) / rvs + 0.5f;
#endif
                        if (pidx.x < lo || pidx.y < lo || pidx.z < lo || pidx.x >= (H.pd[0] - H.buffer) ||
                            pidx.y >= (H.pd[1] - H.buffer) || pidx.z >= (H.pd[2] - H.buffer))
                            continue;
                        pidx = mk(pidx.x + pix * H.pd[0], pidx.y, pidx.z + piy * H.pd[2]);
                        float dc = tex3d(pv, pidx.y, pidx.z, pidx.x);
                        if (dc <= 0.f) continue;
                        col[(size_t)tX + (size_t)bx * ((size_t)tY + (size_t)by * tZ)] = dc;
                    }
        }
        memidx += n;
    }
    return CCCS_OK;
}

}  // namespace

// ================================ extern "C" API =============================================
extern "C" {

int orc_reconfigure(cccs_beam* b, int orientation) {
    // BEAM::reconfigure (beam.cpp:300-331), calc_source_from_angles (:92-97),
    // calc_angles_from_dir (:98-113), closeto (Utilities/math.cpp)
    auto closeto = [](float a, float c) { return fabsf(c - a) <= 1e-6f; };
    const float eps = 1e-6f;
    f3 iso = from_f(b->isocenter);
    if (orientation == 0) {
        b->azimuth = fmodf(b->azimuth, (2 * kPI));
        b->zenith = fmodf(b->zenith, (2 * kPI));
        if (!closeto(b->zenith, 0.f) && (closeto(b->azimuth, 0.f) || closeto(b->azimuth, kPI))) b->zenith = 0.f;
        f3 src = invRotateBeam(mk(0.f, -b->sad, 0.f), b->azimuth, b->zenith, 0.f) + iso;
        f3 dir = normalize(iso - src);
        b->source[0] = src.x; b->source[1] = src.y; b->source[2] = src.z;
        b->direction[0] = dir.x; b->direction[1] = dir.y; b->direction[2] = dir.z;
    } else {
        f3 dir = from_f(b->direction);
        f3 ndir = -1.f * normalize(dir);
        float theta = acosf(-ndir.y);
        float phi = atanf(ndir.z / ndir.x);
        if (closeto(ndir.z, 0.f)) phi = 0.f;
        else if (closeto(ndir.x, 0.f)) phi = kPI / 2.f;
        if (closeto(ndir.x, 0.f) && closeto(ndir.z, 0.f)) theta = (ndir.y <= (-eps) ? 0 : kPI);
        else if (ndir.x < (-eps) || (closeto(ndir.x, 0.f) && ndir.z < (-eps))) theta *= -1;
        b->azimuth = theta;
        b->zenith = phi;
        f3 src = iso - normalize(dir) * b->sad;
        b->source[0] = src.x; b->source[1] = src.y; b->source[2] = src.z;
    }
    return CCCS_OK;
}

int orc_cumulative_kernel(const float* poly_total, int nphi, int ntheta, float* kernel_out, float* theta_out,
                          float* phi_out) {
    // make_cumulative_kernel (kernel.cpp:260-431), total weights only
    const int NR = 24, NT = 48;
    if (ntheta <= 0 || NT % ntheta || ntheta % 2 || nphi <= 0) return CCCS_EINVAL;
    std::vector<float> cum((size_t)ntheta * NR, 0.f);
    for (int tt = 0; tt < ntheta; tt++) {
        int mark = tt * NT / ntheta;
        for (int kk = 0; kk < NT / ntheta; kk++)
            for (int rr = 0; rr < NR; rr++) cum[rr + tt * NR] += poly_total[rr + (mark + kk) * NR];
    }
    for (int tt = 0; tt < ntheta; tt++)
        for (int rr = 1; rr < NR; rr++) cum[rr + tt * NR] += cum[rr - 1 + tt * NR];
    std::memcpy(kernel_out, cum.data(), cum.size() * sizeof(float));
    float ab[49];
    for (int tt = 0; tt < ntheta + 1; tt++) ab[tt] = (180.f / ntheta) * (tt + 1);
    for (int pp = 0; pp < nphi; pp++)
        for (int tt = 0; tt < ntheta / 2; tt++) {
            int aa = pp * ntheta / 2 + tt;
            if (theta_out) theta_out[aa] = 0.5f * (ab[tt] + (tt > 0 ? ab[tt - 1] : 0));
            if (phi_out) phi_out[aa] = (float)((pp + 0.5) * 360. / nphi);
        }
    return CCCS_OK;
}

int orc_kernel_tables(const char* spectrum_file, const char* kernel_dir, const char* radii_file, int nphi,
                      int ntheta, float* kernel_out, float* radii_out, float* theta_out, float* phi_out,
                      float* spec_out, int* nspec, float* poly_total_out, float* cat_out) {
    const int NR = 24, NT = 48, NC = 5;
    // read_spectrum_file (spectra.cpp:7-62)
    FILE* f = fopen(spectrum_file, "r");
    if (!f) return CCCS_EIO;
    float en[30], flu[30], mu[30], mu_en[30];
    char names[30][300];
    int n = 0;
    float sum_fluence = 0;
    while (n < 30 && fscanf(f, "%f %f %f %f %299s", &en[n], &flu[n], &mu[n], &mu_en[n], names[n]) == 5) {
        sum_fluence += flu[n];
        n++;
    }
    fclose(f);
    if (n == 0) return CCCS_EIO;
    for (int i = 0; i < n; i++) flu[i] /= sum_fluence;
    // radial boundaries (read_kernel, kernel.cpp:62-68)
    float radii[NR];
    FILE* rf = fopen(radii_file, "r");
    if (!rf) return CCCS_EIO;
    for (int i = 0; i < NR; i++)
        if (fscanf(rf, "%f", &radii[i]) != 1) { fclose(rf); return CCCS_EIO; }
    fclose(rf);
    // mono kernels (read_kernel, kernel.cpp:80-118)
    std::vector<std::vector<float>> mono((size_t)n * NC, std::vector<float>((size_t)NT * NR));
    for (int e = 0; e < n; e++) {
        std::string path = std::string(kernel_dir) + "/" + names[e];
        FILE* kf = fopen(path.c_str(), "r");
        if (!kf) return CCCS_EIO;
        float unc, mr, ma;
        for (int j = 0; j < NT; j++)
            for (int i = 0; i < NR; i++) {
                for (int c = 0; c < NC; c++)
                    if (fscanf(kf, "%f %f", &mono[(size_t)e * NC + c][i + j * NR], &unc) != 2) { fclose(kf); return CCCS_EIO; }
                if (fscanf(kf, "%f %f", &mr, &unc) != 2 || fscanf(kf, "%f %f", &ma, &unc) != 2) { fclose(kf); return CCCS_EIO; }
            }
        fclose(kf);
    }
    // make_poly_kernel (kernel.cpp:146-211)
    std::vector<float> poly_cat((size_t)NC * NT * NR), poly_tot((size_t)NT * NR);
    for (int j = 0; j < NT; j++)
        for (int i = 0; i < NR; i++) {
            float tot = 0.0f;
            for (int c = 0; c < NC; c++) {
                float v = 0.0f, s = 0.0f;
                for (int e = 0; e < n; e++) {
                    v += flu[e] * en[e] * mu[e] * mono[(size_t)e * NC + c][i + j * NR];
                    s += flu[e] * en[e] * mu[e];
                }
                v /= s;
                poly_cat[(size_t)c * NT * NR + i + j * NR] = v;
                tot += v;
            }
            poly_tot[i + j * NR] = tot;
        }
    if (poly_total_out) std::memcpy(poly_total_out, poly_tot.data(), poly_tot.size() * sizeof(float));
    if (cat_out) {
        // per-category cumulative kernels (same aggregation as the total)
        for (int c = 0; c < NC; c++) {
            std::vector<float> cc((size_t)ntheta * NR, 0.f);
            for (int tt = 0; tt < ntheta; tt++) {
                int mark = tt * NT / ntheta;
                for (int kk = 0; kk < NT / ntheta; kk++)
                    for (int rr = 0; rr < NR; rr++) cc[rr + tt * NR] += poly_cat[(size_t)c * NT * NR + rr + (mark + kk) * NR];
            }
            for (int tt = 0; tt < ntheta; tt++)
                for (int rr = 1; rr < NR; rr++) cc[rr + tt * NR] += cc[rr - 1 + tt * NR];
            std::memcpy(cat_out + (size_t)c * ntheta * NR, cc.data(), cc.size() * sizeof(float));
        }
    }
    int s = orc_cumulative_kernel(poly_tot.data(), nphi, ntheta, kernel_out, theta_out, phi_out);
    if (s) return s;
    if (radii_out) std::memcpy(radii_out, radii, sizeof(radii));
    if (spec_out) {
        for (int e = 0; e < n; e++) {
            spec_out[e] = flu[e];
            spec_out[e + n] = en[e];
            spec_out[e + 2 * n] = mu_en[e];
            spec_out[e + 3 * n] = mu[e];
        }
    }
    if (nspec) *nspec = n;
    return CCCS_OK;
}

int orc_terma(const cccs_problem* p, const cccs_beam* beam, const float* fmap, int beamletnum, float* terma,
              orc_stats* st) {
    if (!problem_ok(p) || !beam || !fmap || !terma) return CCCS_EINVAL;
    terma_bbox(p, *beam, fmap, beamletnum, terma, st);
    return CCCS_OK;
}

int orc_siddon_paths(const cccs_problem* p, const float* source, const float* points, int npoints, int max_segments,
                     int32_t* ijk, float* mid, float* len, int32_t* counts, float* rpl) {
    if (!problem_ok(p) || !source || !points || npoints < 0) return CCCS_EINVAL;
    Prob P(p);
    f3 end = volume_end(P);
    f3 src = from_f(source);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < npoints; i++) {
        int c = 0;
        float r = siddon_rpl(from_f(points + 3 * i), P.start, end, P.voxel, src, P.dens,
                             [&](float fx, float fy, float fz, float l) {
                                 if (c < max_segments) {
                                     size_t o = (size_t)i * max_segments + c;
                                     // voxel whose extent [k-0.5, k+0.5) holds the midpoint
                                     if (ijk) {
                                         ijk[3 * o] = f2i_rd(fx + 0.5f);
                                         ijk[3 * o + 1] = f2i_rd(fy + 0.5f);
                                         ijk[3 * o + 2] = f2i_rd(fz + 0.5f);
                                     }
                                     if (mid) { mid[3 * o] = fx; mid[3 * o + 1] = fy; mid[3 * o + 2] = fz; }
                                     if (len) len[o] = l;
                                 }
                                 c++;
                             },
                             nullptr);
        if (counts) counts[i] = c;
        if (rpl) rpl[i] = r;
    }
    return CCCS_OK;
}

int orc_beam_geometry(const cccs_problem* p, const cccs_beam* beam, const float* terma, cccs_beam_geometry* g) {
    if (!problem_ok(p) || !beam || !terma || !g) return CCCS_EINVAL;
    return beam_geometry_full(p, *beam, terma, g);
}

int orc_beam_cone_stage(const cccs_problem* p, const cccs_beam* beam, const float* fmap, int ray, float* rdens,
                        float* rterma, float* rdose) {
    if (!problem_ok(p) || !beam || !fmap) return CCCS_EINVAL;
    std::vector<float> terma((size_t)p->bbox_size[0] * p->bbox_size[1] * p->bbox_size[2]);
    terma_bbox(p, *beam, fmap, -1, terma.data(), nullptr);
    cccs_beam_geometry g;
    int s = beam_geometry_full(p, *beam, terma.data(), &g);
    if (s != CCCS_OK) return s;
    if (ray < 0 || ray >= g.nrays) return CCCS_EINVAL;
    beam_cone(p, *beam, g, ray, terma.data(), rdens, rterma, rdose, nullptr);
    return CCCS_OK;
}

int orc_beam_dose(const cccs_problem* p, const cccs_beam* beam, const float* fmap, float* dose, orc_stats* st) {
    if (!problem_ok(p) || !beam || !fmap || !dose) return CCCS_EINVAL;
    std::vector<float> terma((size_t)p->bbox_size[0] * p->bbox_size[1] * p->bbox_size[2]);
    terma_bbox(p, *beam, fmap, -1, terma.data(), st);
    cccs_beam_geometry g;
    int s = beam_geometry_full(p, *beam, terma.data(), &g);
    if (s != CCCS_OK) return s;
    for (int ray = 0; ray < g.nrays; ray++) {
        size_t n = (size_t)g.rev_size[ray][0] * g.rev_size[ray][1] * g.rev_size[ray][2];
        std::vector<float> rd(n), rt(n), rdo(n);
        beam_cone(p, *beam, g, ray, terma.data(), rd.data(), rt.data(), rdo.data(), st);
        beam_rev_to_pvcs(p, *beam, g, ray, rdo.data(), dose, st);
    }
    return CCCS_OK;
}

int orc_rev_to_pvcs(const cccs_problem* p, const cccs_beam* beam, const cccs_beam_geometry* g, int ray,
                    const float* rev_dose, float* dose) {
    if (!problem_ok(p) || !beam || !g || !rev_dose || !dose || ray < 0 || ray >= g->nrays) return CCCS_EINVAL;
    beam_rev_to_pvcs(p, *beam, *g, ray, rev_dose, dose, nullptr);
    return CCCS_OK;
}

int orc_row_convolve(const float* dens, const float* terma, int n, const float* kernel, int ntheta, int nradii,
                     const float* radii, int P, float delr, int nphi, float* out) {
    for (int t = 0; t < n; t++)
        out[t] = row_convolve_one(dens, terma, n, t, kernel, ntheta, nradii, radii, (float)P, delr, nphi, nullptr, nullptr);
    return CCCS_OK;
}

int orc_pack_row_convolve(const float* dens, const float* terma, int n, const float* kernel, int ntheta, int nradii,
                          const float* radii, int P, float delr, int nphi, float* out) {
    for (int t = 0; t < n; t++)
        out[t] = pack_row_convolve_one(dens, terma, n, t, kernel, ntheta, nradii, radii, (float)P, delr, nphi, nullptr);
    return CCCS_OK;
}

int64_t orc_sparsify(const float* array, const uint32_t* size, const uint32_t* crop_start, const uint32_t* crop_size,
                     float thresh, uint64_t* keys, float* vals, int64_t cap, float* perc_nonzero,
                     float* perc_dropped) {
    // _sparsify (sparse_io.cpp:22-58)
    uint64_t ntotal = (uint64_t)crop_size[0] * crop_size[1] * crop_size[2];
    if (ntotal == 0) return 0;
    float max_val = *std::max_element(array, array + ntotal);
    uint64_t nnz = 0, ndropped = 0;
    for (uint32_t i = 0; i < crop_size[0]; i++)
        for (uint32_t j = 0; j < crop_size[1]; j++)
            for (uint32_t k = 0; k < crop_size[2]; k++) {
                uint64_t idx = ((uint64_t)k * crop_size[1] + j) * crop_size[0] + i;
                float val = array[idx];
                if (val <= thresh * max_val) {
                    if (val > 0) ndropped++;
                    continue;
                }
                uint64_t true_idx = (((uint64_t)k + crop_start[2]) * size[1] + (j + crop_start[1])) * size[0] + (i + crop_start[0]);
                if ((int64_t)nnz < cap) {
                    keys[nnz] = true_idx;
                    vals[nnz] = val;
                }
                nnz++;
            }
    if (perc_nonzero) *perc_nonzero = (float)(100.0 * nnz / ntotal);
    if (perc_dropped) *perc_dropped = (float)(100.0 * ndropped / ntotal);
    return (int64_t)nnz;
}

int orc_beamlet_anchor(const cccs_problem* p, const cccs_beam* beamp, int idx, float* start3, float* end3,
                       float* biso3, float* angles2) {
    if (!problem_ok(p) || !beamp || !start3 || !end3 || !biso3 || !angles2) return CCCS_EINVAL;
    const cccs_beam& b = *beamp;
    Prob P(p);
    f3 s, e, bi;
    f2 ang;
    f2 bsz{b.beamlet_size[0], b.beamlet_size[1]};
    bool ok = beamlet_anchors(s, e, ang, bi, from_f(b.source), from_f(b.isocenter), (unsigned)idx, bsz, b.fmap_size,
                              P.voxel, P.start, p->bbox_start, p->bbox_size, b.azimuth, b.zenith, b.coll);
    start3[0] = s.x; start3[1] = s.y; start3[2] = s.z;
    end3[0] = e.x; end3[1] = e.y; end3[2] = e.z;
    biso3[0] = bi.x; biso3[1] = bi.y; biso3[2] = bi.z;
    angles2[0] = ang.x; angles2[1] = ang.y;
    return ok ? CCCS_OK : CCCS_EGEOM;
}

int orc_beamlet_dense(const cccs_problem* p, const cccs_beam* beamp, const float* fmap, int mode, float* out,
                      int max_beamlets, int* nactive, orc_stats* st) {
    if (!problem_ok(p) || !beamp || !fmap || mode != CCCS_BEAMLET_FAITHFUL) return CCCS_EINVAL;
    BeamletHooks h;
    return beamlet_run(p, *beamp, fmap, out, max_beamlets, nactive, st, h);
}

int orc_beamlet_cone_stage(const cccs_problem* p, const cccs_beam* beamp, const float* fmap, int ray, float* rev_dens,
                           float* rev_terma, float* rev_dose, uint64_t cap, uint32_t* rev_size) {
    if (!problem_ok(p) || !beamp || !fmap || !rev_size) return CCCS_EINVAL;
    if (ray < 0 || ray >= (int)(p->nphi * (p->ntheta / 2))) return CCCS_EINVAL;
    BeamletHooks h;
    h.stage_ray = ray;
    h.rd = rev_dens;
    h.rt = rev_terma;
    h.rdo = rev_dose;
    h.cap = cap;
    h.rev_size = rev_size;
    int nact = 0;
    return beamlet_run(p, *beamp, fmap, nullptr, 0, &nact, nullptr, h);
}

int orc_beamlet_unpack_linear(const cccs_problem* p, const cccs_beam* beamp, const float* fmap, const float* coef,
                              float* out, int max_beamlets, int* nactive) {
    if (!problem_ok(p) || !beamp || !fmap || !coef || !out) return CCCS_EINVAL;
    BeamletHooks h;
    h.linear = coef;
    return beamlet_run(p, *beamp, fmap, out, max_beamlets, nactive, nullptr, h);
}

// FCBB fluence-map convolution (README.md:15-20; no source in the reference): zero-padded, same
// size, out[y][x] = sum_j sum_i K[j][i] * F[y - (j - kh/2)][x - (i - kw/2)], j outer, i inner.
int orc_fluence_convolve(const float* f, int nmaps, int w, int h, const float* k, int kw, int kh, float* out) {
    if (!f || !k || !out || nmaps < 0 || w < 1 || h < 1 || kw < 1 || kh < 1 || !(kw & 1) || !(kh & 1)) return CCCS_EINVAL;
    const int rx = kw / 2, ry = kh / 2;
    for (int m = 0; m < nmaps; m++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                float acc = 0.f;
                for (int j = 0; j < kh; j++)
                    for (int i = 0; i < kw; i++) {
                        const int yy = y - (j - ry), xx = x - (i - rx);
                        const float v = (xx >= 0 && yy >= 0 && xx < w && yy < h) ? f[((size_t)m * h + yy) * w + xx] : 0.f;
                        acc += k[j * kw + i] * v;
                    }
                out[((size_t)m * h + y) * w + x] = acc;
            }
    return CCCS_OK;
}

// _reduce_by_roi (DoseCalcIO/sparse_io.cpp:62-158): the column scattered into a dense volume, then
// for each ROI the crop box walked ii (x) outer, jj (y), kk (z) inner; every mask voxel advances the
// ROI's element counter and a non-zero dose there is emitted as (mark + element, dose).
int64_t orc_reduce_by_roi(const uint64_t* keys, const float* vals, int64_t nnz, const uint32_t* size, int nroi,
                          const uint8_t* const* masks, const uint32_t* crop_start, const uint32_t* crop_size,
                          uint64_t* out_keys, float* out_vals, int64_t cap, uint64_t* row_block_sizes) {
    const size_t nv = (size_t)size[0] * size[1] * size[2];
    std::vector<float> dense(nv, 0.f);
    for (int64_t i = 0; i < nnz; i++) dense.at(keys[i]) = vals[i];
    int64_t n = 0;
    uint64_t mark = 0;
    for (int r = 0; r < nroi; r++) {
        const uint32_t* c0 = crop_start + 3 * r;
        const uint32_t* cs = crop_size + 3 * r;
        uint64_t kept = 0, elem = 0;
        for (uint32_t ii = 0; ii < cs[0]; ii++)
            for (uint32_t jj = 0; jj < cs[1]; jj++)
                for (uint32_t kk = 0; kk < cs[2]; kk++) {
                    const uint64_t crop_key = ((uint64_t)kk * cs[1] + jj) * cs[0] + ii;
                    const uint64_t orig_key = (((uint64_t)kk + c0[2]) * size[1] + (jj + c0[1])) * size[0] + (ii + c0[0]);
                    if (!masks[r][crop_key]) continue;
                    const float d = dense.at(orig_key);
                    if (d != 0.f) {
                        if (n < cap) {
                            out_keys[n] = mark + elem;
                            out_vals[n] = d;
                        }
                        n++;
                        kept++;
                    }
                    elem++;
                }
        mark += elem;
        row_block_sizes[r] = kept;
    }
    return n;
}

// Preprocess: findFluenceProjection + cudaRaycast (dosecalc-preprocess/cudaMemFunctions.cu:308-393,
// kernels.cuh:136-224).  Nine sub-rays per beamlet (subray index i: offsets (i%3-1, i/3-1) half
// beamlets), bixel centre rotated by inverseRotateBeamAtOriginRHS (geometry.cu:85-92) plus the
// isocentre plus 1e-9, the ray extended to 3x the source-to-centre distance, cudaRaytrace_device
// through the mask (trilinear, border 0, no +0.5 shift).  Beamlet = 1 when any sub-ray's rpl > 0.
int orc_fluence_projection(const cccs_problem* p, const uint8_t* mask, const cccs_beam* beams, int nbeams,
                           float* out) {
    if (!p || !mask || !beams || !out || nbeams < 0) return CCCS_EINVAL;
    const size_t nv = (size_t)p->size[0] * p->size[1] * p->size[2];
    std::vector<float> mf(nv);
    for (size_t i = 0; i < nv; i++) mf[i] = (float)mask[i];
    const Vol v{mf.data(), (int)p->size[0], (int)p->size[1], (int)p->size[2]};
    const f3 start = from_f(p->start), voxel = from_f(p->voxel);
    const f3 end = mk(start.x + (float)p->size[0] * voxel.x, start.y + (float)p->size[1] * voxel.y,
                      start.z + (float)p->size[2] * voxel.z);
    size_t off = 0;
    for (int bi = 0; bi < nbeams; bi++) {
        const cccs_beam& b = beams[bi];
        const int fx = (int)b.fmap_size[0], fy = (int)b.fmap_size[1];
        const f3 src = from_f(b.source), iso = from_f(b.isocenter);
#pragma omp parallel for schedule(dynamic, 8)
        for (int pix = 0; pix < fx * fy; pix++) {
            const int rx = pix % fx, rz = pix / fx;
            bool hit = false;
            for (int sr = 0; sr < 9; sr++) {
                const int sx = sr % 3 - 1, sz = sr / 3 - 1;
                f3 fcs = mk(b.beamlet_size[0] * ((float)rx + 0.5f * (float)sx - (float)fx * 0.5f + 0.5f), 0.f,
                            b.beamlet_size[1] * ((float)rz + 0.5f * (float)sz - (float)fy * 0.5f + 0.5f));
                f3 ctr = invRotateBeam(fcs, b.azimuth, b.zenith, b.coll) + iso;
                ctr = ctr + 1e-9f;
                const f3 sink = src + 3.f * (ctr - src);
                if (siddon_rpl(sink, start, end, voxel, src, v, [](float, float, float, float) {}, nullptr) > 0.f)
                    hit = true;
            }
            out[off + (size_t)pix] = hit ? 1.f : 0.f;
        }
        off += (size_t)fx * fy;
    }
    return CCCS_OK;
}

}  // extern "C"


// ---- DICOM ingestion ----------------------------------------------------------------------------
// Restated from dosecalc-preprocess (cudaMemFunctions.cu:83-294, kernels.cuh:13-135) and
// DoseCalcIO/roi.cpp:245-345, one output element at a time.
namespace {
float orc_tex_weight(float f) { return std::floor(f * 256.f + 0.5f) * (1.f / 256.f); }
float orc_ct_at(const float* hu, const uint32_t* n, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= (int)n[0] || j >= (int)n[1] || k >= (int)n[2]) return 0.f;
    return hu[(size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * k)];
}
void orc_line(uint8_t* img, int w, int h, int x0, int y0, int x1, int y1) {
    // 8-connected line, error form of Bresenham's algorithm
    int dx = x1 > x0 ? x1 - x0 : x0 - x1, dy = y1 > y0 ? y1 - y0 : y0 - y1;
    int sx = x0 < x1 ? 1 : -1, sy = y0 < y1 ? 1 : -1, err = dx - dy;
    while (true) {
        if (x0 >= 0 && x0 < w && y0 >= 0 && y0 < h) img[(size_t)y0 * w + x0] = 1;
        if (x0 == x1 && y0 == y1) return;
        int e2 = 2 * err;
        if (e2 > -dy) { err -= dy; x0 += sx; }
        if (e2 < dx) { err += dx; y0 += sy; }
    }
}
}  // namespace

extern "C" int orc_ct_to_density(const float* hu, const uint32_t* size, const float* vox, float iso,
                                 const float* lut_hu, const float* lut_dens, int nlut, float* out, uint32_t* out_size) {
    for (int a = 0; a < 3; a++) out_size[a] = (uint32_t)std::ceil(vox[a] * ((float)size[a] / iso));
    if (!out) return 0;
    for (uint32_t Z = 0; Z < out_size[2]; Z++)
        for (uint32_t Y = 0; Y < out_size[1]; Y++)
            for (uint32_t X = 0; X < out_size[0]; X++) {
                const float p[3] = {((float)X * iso) / vox[0], ((float)Y * iso) / vox[1], ((float)Z * iso) / vox[2]};
                float b[3], w[3];
                int ix[3];
                for (int a = 0; a < 3; a++) {
                    b[a] = (p[a] + 0.5f) - 0.5f;
                    const float f = std::floor(b[a]);
                    ix[a] = (int)f;
                    w[a] = orc_tex_weight(b[a] - f);
                }
                float c[2][2];
                for (int dz = 0; dz < 2; dz++)
                    for (int dy = 0; dy < 2; dy++)
                        c[dz][dy] = (1.f - w[0]) * orc_ct_at(hu, size, ix[0], ix[1] + dy, ix[2] + dz) +
                                    w[0] * orc_ct_at(hu, size, ix[0] + 1, ix[1] + dy, ix[2] + dz);
                const float c0 = (1.f - w[1]) * c[0][0] + w[1] * c[0][1];
                const float c1 = (1.f - w[1]) * c[1][0] + w[1] * c[1][1];
                float v = (1.f - w[2]) * c0 + w[2] * c1;
                if (nlut >= 2) {
                    if (v <= lut_hu[0]) v = lut_dens[0];
                    else if (v >= lut_hu[nlut - 1])
                        v = lut_dens[nlut - 2] + (v - lut_hu[nlut - 2]) * (lut_dens[nlut - 1] - lut_dens[nlut - 2]) /
                                                     (lut_hu[nlut - 1] - lut_hu[nlut - 2]);
                    else {
                        int lo = 0;
                        while (lut_hu[lo] < v && lo < nlut - 1) lo++;
                        lo--;
                        v = lut_dens[lo] + (v - lut_hu[lo]) * (lut_dens[lo + 1] - lut_dens[lo]) / (lut_hu[lo + 1] - lut_hu[lo]);
                    }
                } else {
                    v = (v + 1024.0f) / 1024.0f;
                    if (v < 0.05f) v = 0.05f;
                }
                out[(size_t)X + (size_t)out_size[0] * ((size_t)Y + (size_t)out_size[1] * Z)] = v;
            }
    return 0;
}

extern "C" int orc_mask_resample(const uint8_t* in, const uint32_t* is, const float* ist, const float* isp,
                                 const uint32_t* os, const float* ost, const float* osp, uint8_t* out) {
    for (uint32_t Z = 0; Z < os[2]; Z++)
        for (uint32_t Y = 0; Y < os[1]; Y++)
            for (uint32_t X = 0; X < os[0]; X++) {
                const uint32_t o[3] = {X, Y, Z};
                int t[3];
                bool inside = true;
                for (int a = 0; a < 3; a++) {
                    const float u = ((float)o[a] * osp[a] + ost[a] - ist[a]) / isp[a];
                    t[a] = (int)std::floor(u + 0.5f);
                    inside = inside && t[a] >= 0 && t[a] < (int)is[a];
                }
                const uint8_t v = inside ? in[(size_t)t[0] + (size_t)is[0] * ((size_t)t[1] + (size_t)is[1] * t[2])] : 0;
                out[(size_t)X + (size_t)os[0] * ((size_t)Y + (size_t)os[1] * Z)] = v >= 1 ? 1 : 0;
            }
    return 0;
}

extern "C" int orc_roi_extents(const float* pts, uint64_t n, const uint32_t* size, const float* start, const float* sp,
                               uint32_t* cs, uint32_t* cz) {
    for (int a = 0; a < 3; a++) {
        float mn = std::numeric_limits<float>::max(), mx = std::numeric_limits<float>::min();
        for (uint64_t i = 0; i < n; i++) {
            mn = pts[3 * i + a] < mn ? pts[3 * i + a] : mn;
            mx = pts[3 * i + a] > mx ? pts[3 * i + a] : mx;
        }
        int s = (int)std::floor((0.1f * mn - start[a]) / sp[a]), e = (int)std::ceil((0.1f * mx - start[a]) / sp[a]);
        s = s < 0 ? 0 : (s > (int)size[a] ? (int)size[a] : s);
        e = e < 0 ? 0 : (e > (int)size[a] ? (int)size[a] : e);
        cs[a] = (uint32_t)s;
        cz[a] = (uint32_t)(e - s);
    }
    return 0;
}

extern "C" int orc_roi_centroid(const uint8_t* mask, const uint32_t* size, const float* start, const float* sp,
                                float* c) {
    float sx = 0.f, sy = 0.f, sz = 0.f, cnt = 0.f;
    for (uint32_t Z = 0; Z < size[2]; Z++)
        for (uint32_t Y = 0; Y < size[1]; Y++)
            for (uint32_t X = 0; X < size[0]; X++)
                if (mask[(size_t)X + (size_t)size[0] * ((size_t)Y + (size_t)size[1] * Z)]) {
                    sx += start[0] + ((float)X + 0.5f) * sp[0];
                    sy += start[1] + ((float)Y + 0.5f) * sp[1];
                    sz += start[2] + ((float)Z + 0.5f) * sp[2];
                    cnt += 1.f;
                }
    c[0] = sx / cnt;
    c[1] = sy / cnt;
    c[2] = sz / cnt;
    return 0;
}

extern "C" int orc_contour_rasterize(const float* pts, const uint32_t* counts, int nsub, const uint32_t* size,
                                     const float* start, const float* sp, uint8_t* mask) {
    const int w = (int)size[0], h = (int)size[1];
    const size_t plane = (size_t)w * h;
    std::fill(mask, mask + plane * size[2], (uint8_t)0);
    std::vector<uint8_t> img(plane);
    size_t mark = 0;
    for (int c = 0; c < nsub; c++) {
        const int np = (int)counts[c];
        if (!np) continue;
        const int z = (int)std::round((0.1f * pts[3 * mark + 2] - start[2]) / sp[2]);
        std::vector<int> xs(np), ys(np);
        for (int q = 0; q < np; q++) {
            xs[q] = (int)std::floor((0.1f * pts[3 * (mark + q)] - start[0]) / sp[0]);
            ys[q] = (int)std::floor((0.1f * pts[3 * (mark + q) + 1] - start[1]) / sp[1]);
        }
        mark += (size_t)np;
        if (z < 0 || z >= (int)size[2]) continue;
        std::fill(img.begin(), img.end(), (uint8_t)0);
        for (int q = 0; q < np; q++) orc_line(img.data(), w, h, xs[q], ys[q], xs[(q + 1) % np], ys[(q + 1) % np]);
        if (np >= 3)
            for (int y = 0; y < h; y++) {
                // crossings x = x0 + (y - y0)(x1 - x0)/(y1 - y0) of edges with y0 <= y < y1, as doubles
                // compared exactly through their rational form
                std::vector<std::pair<long long, long long>> xc;
                for (int q = 0; q < np; q++) {
                    int x0 = xs[q], y0 = ys[q], x1 = xs[(q + 1) % np], y1 = ys[(q + 1) % np];
                    if (y0 == y1) continue;
                    if (y0 > y1) { std::swap(x0, x1); std::swap(y0, y1); }
                    if (y < y0 || y >= y1) continue;
                    xc.emplace_back((long long)x0 * (y1 - y0) + (long long)(y - y0) * (x1 - x0), (long long)(y1 - y0));
                }
                std::sort(xc.begin(), xc.end(), [](const std::pair<long long, long long>& a,
                                                   const std::pair<long long, long long>& b) {
                    return (__int128)a.first * b.second < (__int128)b.first * a.second;
                });
                for (size_t k = 0; k + 1 < xc.size(); k += 2) {
                    const long long lo = (long long)std::ceil((long double)xc[k].first / xc[k].second);
                    const long long hi = (long long)std::floor((long double)xc[k + 1].first / xc[k + 1].second);
                    for (long long x = lo < 0 ? 0 : lo; x <= hi && x < w; x++) img[(size_t)y * w + x] = 1;
                }
            }
        uint8_t* dst = mask + plane * z;
        for (size_t k = 0; k < plane; k++) dst[k] = (uint8_t)((img[k] + dst[k]) > 0);
    }
    return 0;
}
