// host.cpp -- host-side input preparation for the CCCS path: rotation matrices, beam
// orientation (BEAM::reconfigure), and the poly-energetic / cumulative kernel tables.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cccs_internal.hpp"
#include "errors.hpp"

namespace cccs {

// float PI of the reference (CUDART_PI_F, include/dosecalc_defs.h:8)
static const float kPIf = 3.14159265358979323846f;

dm33 dm_axis(double x, double y, double z, double t) {
    // Rodrigues: R = c I + s [r]x + (1-c) r r^T   (right-hand rotation by t about r)
    double s = std::sin(t), c = std::cos(t), k = 1.0 - c;
    dm33 r;
    r.m[0] = c + k * x * x;     r.m[1] = k * x * y - s * z; r.m[2] = k * x * z + s * y;
    r.m[3] = k * y * x + s * z; r.m[4] = c + k * y * y;     r.m[5] = k * y * z - s * x;
    r.m[6] = k * z * x - s * y; r.m[7] = k * z * y + s * x; r.m[8] = c + k * z * z;
    return r;
}
dm33 dm_mul(const dm33& a, const dm33& b) {
    dm33 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += a.m[3 * i + k] * b.m[3 * k + j];
            r.m[3 * i + j] = s;
        }
    return r;
}
dm33 dm_transpose(const dm33& a) {
    dm33 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[3 * i + j] = a.m[3 * j + i];
    return r;
}
m33 to_f(const dm33& a) {
    m33 r;
    for (int i = 0; i < 9; i++) r.m[i] = (float)a.m[i];
    return r;
}
v3 dm_apply_f(const dm33& a, v3 v) {
    double x = v.x, y = v.y, z = v.z;
    return mk3((float)(a.m[0] * x + a.m[1] * y + a.m[2] * z), (float)(a.m[3] * x + a.m[4] * y + a.m[5] * z),
               (float)(a.m[6] * x + a.m[7] * y + a.m[8] * z));
}

// BEV <- PVCS: rotate about the couch axis (sin(-phi), 0, cos(-phi)) by -theta, then about y by
// phi + coll (geometry.cu:69-75).  The angles are the FP32 values the beam carries.
dm33 dm_beam_fwd(float theta, float phi, float coll) {
    double th = theta, ph = phi;
    dm33 couch = dm_axis(std::sin(-ph), 0.0, std::cos(-ph), -th);
    dm33 roll = dm_axis(0.0, 1.0, 0.0, (double)(phi + coll));
    return dm_mul(roll, couch);
}
dm33 dm_beam_inv(float theta, float phi, float coll) { return dm_transpose(dm_beam_fwd(theta, phi, coll)); }
// REV <- BEV: about (cos phi, 0, -sin phi) by -theta, then about y by -phi (geometry.cu:126-131)
dm33 dm_kernel_fwd(float theta, float phi) {
    double th = theta, ph = phi;
    dm33 tilt = dm_axis(std::cos(ph), 0.0, -std::sin(ph), -th);
    dm33 roll = dm_axis(0.0, 1.0, 0.0, -ph);
    return dm_mul(roll, tilt);
}
dm33 dm_kernel_inv(float theta, float phi) { return dm_transpose(dm_kernel_fwd(theta, phi)); }
// beamlet bev -> BEV is a rotation about (sin(-phi), 0, cos(-phi)) by -theta (geometry.cu:110-115)
dm33 dm_beamlet_inv(float theta, float phi) {
    double th = theta, ph = phi;
    return dm_axis(std::sin(-ph), 0.0, std::cos(-ph), -th);
}
dm33 dm_beamlet_fwd(float theta, float phi) { return dm_transpose(dm_beamlet_inv(theta, phi)); }

static Rot2 make_rot2(v3 a1, float t1, v3 a2, float t2) {
    Rot2 r;
    r.a1 = a1;
    r.a2 = a2;
    sincosf(t1, &r.s1, &r.c1);
    sincosf(t2, &r.s2, &r.c2);
    return r;
}
Rot2 rot2_beam_fwd(float theta, float phi, float coll) {
    float s, c;
    sincosf(-phi, &s, &c);
    return make_rot2(mk3(s, 0.f, c), -theta, mk3(0.f, 1.f, 0.f), phi + coll);
}
Rot2 rot2_beam_inv(float theta, float phi, float coll) {
    float s, c;
    sincosf(-phi, &s, &c);
    return make_rot2(mk3(0.f, 1.f, 0.f), -(phi + coll), mk3(s, 0.f, c), theta);
}
Rot2 rot2_kernel_fwd(float theta, float phi) {
    float s, c;
    sincosf(phi, &s, &c);
    return make_rot2(mk3(c, 0.f, -s), -theta, mk3(0.f, 1.f, 0.f), -phi);
}
Rot2 rot2_kernel_inv(float theta, float phi) {
    float s, c;
    sincosf(phi, &s, &c);
    return make_rot2(mk3(0.f, 1.f, 0.f), phi, mk3(c, 0.f, -s), theta);
}

namespace {

// make_cumulative_kernel (DoseCalcIO/kernel.cpp:260-431): aggregate the 48 polar bins of the
// poly kernel into ntheta cones, cumulative-sum along radius, and derive the cone angles.
int cumulative(const float* poly_total, int nphi, int ntheta, float* kernel_out, float* theta_out, float* phi_out) {
    const int NR = 24, NT = 48;
    if (!poly_total || !kernel_out || ntheta <= 0 || ntheta > NT || NT % ntheta != 0 || ntheta % 2 != 0 || nphi <= 0)
        return set_error(CCCS_EINVAL, "ntheta must be even and divide 48; nphi must be positive");
    std::vector<float> k((size_t)ntheta * NR, 0.f);
    const int group = NT / ntheta;
    for (int t = 0; t < ntheta; t++)
        for (int g = 0; g < group; g++)
            for (int r = 0; r < NR; r++) k[(size_t)t * NR + r] += poly_total[(size_t)(t * group + g) * NR + r];
    for (int t = 0; t < ntheta; t++)
        for (int r = 1; r < NR; r++) k[(size_t)t * NR + r] += k[(size_t)t * NR + r - 1];
    std::copy(k.begin(), k.end(), kernel_out);
    // boundaries (180/ntheta)*(t+1); cone polar angle = mid-point of its bin; azimuth (p+1/2)*360/nphi
    std::vector<float> bound(ntheta + 1);
    for (int t = 0; t <= ntheta; t++) bound[t] = (180.f / ntheta) * (t + 1);
    for (int p = 0; p < nphi; p++)
        for (int t = 0; t < ntheta / 2; t++) {
            int a = p * ntheta / 2 + t;
            if (theta_out) theta_out[a] = 0.5f * (bound[t] + (t > 0 ? bound[t - 1] : 0.f));
            if (phi_out) phi_out[a] = (float)((p + 0.5) * 360.0 / nphi);
        }
    return CCCS_OK;
}

}  // namespace
}  // namespace cccs

using namespace cccs;

extern "C" int cccs_cumulative_kernel(const float* poly_total, int nphi, int ntheta, float* kernel_out,
                                      float* theta_out, float* phi_out) {
    clear_error();
    return cumulative(poly_total, nphi, ntheta, kernel_out, theta_out, phi_out);
}

// read_spectrum_file (spectra.cpp:7-62), read_kernel (kernel.cpp:27-122) and make_poly_kernel
// (kernel.cpp:123-258), then the cumulative table.
extern "C" int cccs_poly_kernel_from_files(const char* spectrum_file, const char* kernel_dir, const char* radii_file,
                                           float* poly_out, float* radii_out, float* spec_out, int* nspec) {
    clear_error();
    const int NR = 24, NT = 48, NC = 5;
    if (!spectrum_file || !kernel_dir || !radii_file) return set_error(CCCS_EINVAL, "null path");
    FILE* f = std::fopen(spectrum_file, "r");
    if (!f) return set_error(CCCS_EIO, std::string("cannot open spectrum ") + spectrum_file);
    struct Bin { float e, flu, mu, mu_en; char name[300]; };
    std::vector<Bin> bins;
    Bin b;
    while (bins.size() < (size_t)kMaxSpec && std::fscanf(f, "%f %f %f %f %299s", &b.e, &b.flu, &b.mu, &b.mu_en, b.name) == 5)
        bins.push_back(b);
    std::fclose(f);
    if (bins.empty()) return set_error(CCCS_EIO, "empty spectrum");
    float total = 0.f;
    for (auto& x : bins) total += x.flu;
    for (auto& x : bins) x.flu /= total;
    float radii[NR];
    FILE* rf = std::fopen(radii_file, "r");
    if (!rf) return set_error(CCCS_EIO, std::string("cannot open radii ") + radii_file);
    for (int i = 0; i < NR; i++)
        if (std::fscanf(rf, "%f", &radii[i]) != 1) { std::fclose(rf); return set_error(CCCS_EIO, "short radii file"); }
    std::fclose(rf);
    const size_t plane = (size_t)NT * NR;
    std::vector<float> mono(bins.size() * NC * plane);
    for (size_t e = 0; e < bins.size(); e++) {
        std::string path = std::string(kernel_dir) + "/" + bins[e].name;
        FILE* kf = std::fopen(path.c_str(), "r");
        if (!kf) return set_error(CCCS_EIO, "cannot open kernel " + path);
        // per (theta, radius): 5 category (value, uncertainty) pairs, then mean radius and angle
        for (int j = 0; j < NT; j++)
            for (int i = 0; i < NR; i++) {
                float v, u;
                for (int c = 0; c < NC; c++) {
                    if (std::fscanf(kf, "%f %f", &v, &u) != 2) { std::fclose(kf); return set_error(CCCS_EIO, "short kernel " + path); }
                    mono[(e * NC + c) * plane + (size_t)j * NR + i] = v;
                }
                for (int extra = 0; extra < 2; extra++)
                    if (std::fscanf(kf, "%f %f", &v, &u) != 2) { std::fclose(kf); return set_error(CCCS_EIO, "short kernel " + path); }
            }
        std::fclose(kf);
    }
    // poly kernel: per category, fluence*energy*mu weighted mean over energies; total = sum
    std::vector<float> poly(plane);
    for (int j = 0; j < NT; j++)
        for (int i = 0; i < NR; i++) {
            float tot = 0.0f;
            for (int c = 0; c < NC; c++) {
                float acc = 0.0f, wsum = 0.0f;
                for (size_t e = 0; e < bins.size(); e++) {
                    float w = bins[e].flu * bins[e].e * bins[e].mu;
                    acc += w * mono[(e * NC + c) * plane + (size_t)j * NR + i];
                    wsum += w;
                }
                acc /= wsum;
                tot += acc;
            }
            poly[(size_t)j * NR + i] = tot;
        }
    if (poly_out) std::copy(poly.begin(), poly.end(), poly_out);
    if (radii_out) std::memcpy(radii_out, radii, sizeof(radii));
    const int n = (int)bins.size();
    if (spec_out)
        for (int e = 0; e < n; e++) {
            spec_out[e] = bins[e].flu;
            spec_out[n + e] = bins[e].e;
            spec_out[2 * n + e] = bins[e].mu_en;
            spec_out[3 * n + e] = bins[e].mu;
        }
    if (nspec) *nspec = n;
    return CCCS_OK;
}

// BEAM::reconfigure (DoseCalcIO/beam.cpp:300-331) with calc_source_from_angles (:92-97) and
// calc_angles_from_dir (:98-113).
extern "C" int cccs_beam_reconfigure(cccs_beam* beam, int orientation) {
    clear_error();
    if (!beam) return set_error(CCCS_EINVAL, "null beam");
    auto close = [](float a, float b) { return std::fabs(b - a) <= 1e-6f; };
    const float eps = 1e-6f;
    v3 iso = ld3(beam->isocenter);
    if (orientation == 0) {
        beam->azimuth = std::fmod(beam->azimuth, 2 * kPIf);
        beam->zenith = std::fmod(beam->zenith, 2 * kPIf);
        if (!close(beam->zenith, 0.f) && (close(beam->azimuth, 0.f) || close(beam->azimuth, kPIf))) beam->zenith = 0.f;
        // source = iso + inverse beam rotation of (0, -sad, 0), in the reference's float Rodrigues
        // form (calc_source_from_angles, beam.cpp:92-97) so the source matches it bit for bit
        v3 s = rot2(rot2_beam_inv(beam->azimuth, beam->zenith, 0.f), mk3(0.f, -beam->sad, 0.f)) + iso;
        v3 d = norm3(iso - s);
        beam->source[0] = s.x; beam->source[1] = s.y; beam->source[2] = s.z;
        beam->direction[0] = d.x; beam->direction[1] = d.y; beam->direction[2] = d.z;
    } else if (orientation == 1) {
        v3 dir = ld3(beam->direction);
        v3 nd = -1.f * norm3(dir);
        float theta = std::acos(-nd.y);
        float phi = std::atan(nd.z / nd.x);
        if (close(nd.z, 0.f)) phi = 0.f;
        else if (close(nd.x, 0.f)) phi = kPIf / 2.f;
        if (close(nd.x, 0.f) && close(nd.z, 0.f)) theta = (nd.y <= -eps ? 0.f : kPIf);
        else if (nd.x < -eps || (close(nd.x, 0.f) && nd.z < -eps)) theta *= -1;
        beam->azimuth = theta;
        beam->zenith = phi;
        v3 s = iso - norm3(dir) * beam->sad;
        beam->source[0] = s.x; beam->source[1] = s.y; beam->source[2] = s.z;
    } else {
        return set_error(CCCS_EINVAL, "orientation must be 0 (angles) or 1 (direction)");
    }
    return CCCS_OK;
}

extern "C" int cccs_build_kernel_from_files(const char* spectrum_file, const char* kernel_dir, const char* radii_file,
                                            int nphi, int ntheta, float* kernel_out, float* radii_out,
                                            float* theta_out, float* phi_out, float* spec_out, int* nspec) {
    std::vector<float> poly(48 * 24);
    int s = cccs_poly_kernel_from_files(spectrum_file, kernel_dir, radii_file, poly.data(), radii_out, spec_out, nspec);
    if (s != CCCS_OK) return s;
    return cumulative(poly.data(), nphi, ntheta, kernel_out, theta_out, phi_out);
}
