// ingest.hip -- the device side of DICOM ingestion: the CT volume resampled onto the isotropic
// calculation grid with the HU -> density lookup (cudaCreateTexIso + cudaMakeIsotropicWithLUT /
// cudaMakeIsotropic + deviceHU2dens, CCCS/dosecalc-preprocess/cudaMemFunctions.cu:83-216,
// kernels.cuh:13-135), and structure masks resampled from the CT grid onto the calculation grid
// and thresholded (cudaResample + cudaThreshold, CCCS/CudaUtilities/array_sum.cuh:13-58, as
// generateContourMask calls them, cudaMemFunctions.cu:257-294).
//
// The reference samples 3-D textures: linear filtering with border addressing for the CT and
// point sampling for the mask.  Here both are explicit gathers from a linear array in HBM: the
// linear fetch computes the texel-space position x - 0.5, its floor and fraction, quantises the
// fraction to 8 fractional bits (the texture unit's filtering weights) and interpolates the eight
// neighbours (zero outside the volume); the point fetch takes texel floor(x).  One lane per
// output voxel, x fastest (coalesced stores).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "context.hpp"
#include "device_common.hpp"

namespace cccs {
namespace {

struct IsoArgs {
    const float* ct;
    int nx, ny, nz;           // CT grid
    float vx, vy, vz;         // CT voxel size [cm]
    float iso;                // output voxel size [cm]
    int ox, oy, oz;           // output grid
    const float* lut_hu;      // sorted HU points, or nullptr: (HU + 1024) / 1024 clamped at 0.05
    const float* lut_dens;
    int nlut;
    float* out;
};

// texture-unit style linear filtering weight: fraction in 1/256 steps
__host__ __device__ __forceinline__ float tex_weight(float frac) { return floorf(frac * 256.f + 0.5f) * (1.f / 256.f); }

__device__ __forceinline__ float ct_at(const IsoArgs& a, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= a.nx || j >= a.ny || k >= a.nz) return 0.f;
    return a.ct[(size_t)i + (size_t)a.nx * ((size_t)j + (size_t)a.ny * k)];
}

__global__ void __launch_bounds__(256) iso_density_kernel(IsoArgs a) {
    const int X = blockIdx.x * 64 + threadIdx.x, Y = blockIdx.y * 4 + threadIdx.y, Z = blockIdx.z;
    if (X >= a.ox || Y >= a.oy || Z >= a.oz) return;
    // physical position in the output grid, then CT texel coordinates (kernels.cuh:99-111)
    const float px = ((float)X * a.iso) / a.vx, py = ((float)Y * a.iso) / a.vy, pz = ((float)Z * a.iso) / a.vz;
    const float bx = (px + 0.5f) - 0.5f, by = (py + 0.5f) - 0.5f, bz = (pz + 0.5f) - 0.5f;
    const float fx = floorf(bx), fy = floorf(by), fz = floorf(bz);
    const int i = (int)fx, j = (int)fy, k = (int)fz;
    const float ax = tex_weight(bx - fx), ay = tex_weight(by - fy), az = tex_weight(bz - fz);
    const float c00 = (1.f - ax) * ct_at(a, i, j, k) + ax * ct_at(a, i + 1, j, k);
    const float c10 = (1.f - ax) * ct_at(a, i, j + 1, k) + ax * ct_at(a, i + 1, j + 1, k);
    const float c01 = (1.f - ax) * ct_at(a, i, j, k + 1) + ax * ct_at(a, i + 1, j, k + 1);
    const float c11 = (1.f - ax) * ct_at(a, i, j + 1, k + 1) + ax * ct_at(a, i + 1, j + 1, k + 1);
    const float c0 = (1.f - ay) * c00 + ay * c10, c1 = (1.f - ay) * c01 + ay * c11;
    float value = (1.f - az) * c0 + az * c1;
    if (a.lut_hu) {
        // clamp below, extrapolate above from the last two points, else interpolate (kernels.cuh:113-129)
        const int n = a.nlut;
        if (value <= a.lut_hu[0]) {
            value = a.lut_dens[0];
        } else if (value >= a.lut_hu[n - 1]) {
            value = a.lut_dens[n - 2] + (value - a.lut_hu[n - 2]) * (a.lut_dens[n - 1] - a.lut_dens[n - 2]) /
                                            (a.lut_hu[n - 1] - a.lut_hu[n - 2]);
        } else {
            int lo = 0;
            while (a.lut_hu[lo] < value && lo < n - 1) lo++;
            lo--;
            value = a.lut_dens[lo] + (value - a.lut_hu[lo]) * (a.lut_dens[lo + 1] - a.lut_dens[lo]) /
                                         (a.lut_hu[lo + 1] - a.lut_hu[lo]);
        }
    } else {
        // deviceHU2dens (kernels.cuh:16-29)
        value = (value + 1024.0f) / 1024.0f;
        if (value < 0.05f) value = 0.05f;
    }
    a.out[(size_t)X + (size_t)a.ox * ((size_t)Y + (size_t)a.oy * Z)] = value;
}

struct MaskArgs {
    const uint8_t* in;
    int ix, iy, iz;
    float is_x, is_y, is_z, isp_x, isp_y, isp_z;  // input start / spacing [cm]
    int ox, oy, oz;
    float os_x, os_y, os_z, osp_x, osp_y, osp_z;  // output start / spacing [cm]
    uint8_t* out;
};

__global__ void __launch_bounds__(256) mask_resample_kernel(MaskArgs a) {
    const int X = blockIdx.x * 64 + threadIdx.x, Y = blockIdx.y * 4 + threadIdx.y, Z = blockIdx.z;
    if (X >= a.ox || Y >= a.oy || Z >= a.oz) return;
    const float ux = ((float)X * a.osp_x + a.os_x - a.is_x) / a.isp_x;
    const float uy = ((float)Y * a.osp_y + a.os_y - a.is_y) / a.isp_y;
    const float uz = ((float)Z * a.osp_z + a.os_z - a.is_z) / a.isp_z;
    const int i = (int)floorf(ux + 0.5f), j = (int)floorf(uy + 0.5f), k = (int)floorf(uz + 0.5f);
    uint8_t v = 0;
    if (i >= 0 && j >= 0 && k >= 0 && i < a.ix && j < a.iy && k < a.iz)
        v = a.in[(size_t)i + (size_t)a.ix * ((size_t)j + (size_t)a.iy * k)];
    a.out[(size_t)X + (size_t)a.ox * ((size_t)Y + (size_t)a.oy * Z)] = v >= 1 ? 1 : 0;
}

struct DevMem {
    void* p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return set_error(CCCS_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
        return CCCS_OK;
    }
};

}  // namespace
}  // namespace cccs

using namespace cccs;

extern "C" int cccs_ct_to_density(int device, const float* hu, const uint32_t size[3], const float voxel_cm[3],
                                  float iso_voxel_cm, const float* lut_hu, const float* lut_dens, int nlut,
                                  float* density, uint64_t cap, uint32_t out_size[3]) {
    clear_error();
    if (!size || !voxel_cm || !out_size) return set_error(CCCS_EINVAL, "null argument");
    if (!(iso_voxel_cm > 0.f)) return set_error(CCCS_EINVAL, "iso voxel size must be > 0");
    for (int a = 0; a < 3; a++)
        if (!size[a] || !(voxel_cm[a] > 0.f)) return set_error(CCCS_EINVAL, "empty CT grid or non-positive voxel size");
    if (nlut < 0 || nlut == 1 || (nlut > 0 && (!lut_hu || !lut_dens)))
        return set_error(CCCS_EINVAL, "the lookup table needs at least two points");
    // output grid: ceil(voxel * (size / iso)) per axis (cudaMemFunctions.cu:134)
    for (int a = 0; a < 3; a++) out_size[a] = (uint32_t)std::ceil(voxel_cm[a] * ((float)size[a] / iso_voxel_cm));
    if (!density) return CCCS_OK;
    if (!hu) return set_error(CCCS_EINVAL, "null CT volume");
    const uint64_t nin = (uint64_t)size[0] * size[1] * size[2];
    const uint64_t nout = (uint64_t)out_size[0] * out_size[1] * out_size[2];
    if (cap < nout) return set_error(CCCS_EINVAL, "output buffer too small");
    if (nout >= (uint64_t(1) << 31) || nin >= (uint64_t(1) << 31)) return set_error(CCCS_EINVAL, "grid too large");
    CCCS_HIP(hipSetDevice(device));
    DevMem dct, dout, dlut;
    CCCS_TRY(dct.alloc(nin * sizeof(float)));
    CCCS_TRY(dout.alloc(nout * sizeof(float)));
    CCCS_TRY(dlut.alloc(2 * sizeof(float) * (size_t)std::max(nlut, 1)));
    CCCS_HIP(hipMemcpy(dct.p, hu, nin * sizeof(float), hipMemcpyHostToDevice));
    IsoArgs a;
    a.ct = static_cast<const float*>(dct.p);
    a.nx = (int)size[0]; a.ny = (int)size[1]; a.nz = (int)size[2];
    a.vx = voxel_cm[0]; a.vy = voxel_cm[1]; a.vz = voxel_cm[2];
    a.iso = iso_voxel_cm;
    a.ox = (int)out_size[0]; a.oy = (int)out_size[1]; a.oz = (int)out_size[2];
    a.nlut = nlut;
    a.lut_hu = a.lut_dens = nullptr;
    if (nlut > 0) {
        float* l = static_cast<float*>(dlut.p);
        CCCS_HIP(hipMemcpy(l, lut_hu, nlut * sizeof(float), hipMemcpyHostToDevice));
        CCCS_HIP(hipMemcpy(l + nlut, lut_dens, nlut * sizeof(float), hipMemcpyHostToDevice));
        a.lut_hu = l;
        a.lut_dens = l + nlut;
    }
    a.out = static_cast<float*>(dout.p);
    dim3 blk(64, 4, 1), grd((a.ox + 63) / 64, (a.oy + 3) / 4, a.oz);
    hipLaunchKernelGGL(iso_density_kernel, grd, blk, 0, 0, a);
    CCCS_HIP(hipGetLastError());
    CCCS_HIP(hipMemcpy(density, dout.p, nout * sizeof(float), hipMemcpyDeviceToHost));
    return CCCS_OK;
}

extern "C" int cccs_mask_resample(int device, const uint8_t* mask, const uint32_t in_size[3],
                                  const float in_start_cm[3], const float in_spacing_cm[3], const uint32_t out_size[3],
                                  const float out_start_cm[3], const float out_spacing_cm[3], uint8_t* out) {
    clear_error();
    if (!mask || !in_size || !in_start_cm || !in_spacing_cm || !out_size || !out_start_cm || !out_spacing_cm || !out)
        return set_error(CCCS_EINVAL, "null argument");
    for (int a = 0; a < 3; a++)
        if (!in_size[a] || !out_size[a] || !(in_spacing_cm[a] > 0.f))
            return set_error(CCCS_EINVAL, "empty grid or non-positive spacing");
    const uint64_t nin = (uint64_t)in_size[0] * in_size[1] * in_size[2];
    const uint64_t nout = (uint64_t)out_size[0] * out_size[1] * out_size[2];
    if (nout >= (uint64_t(1) << 31) || nin >= (uint64_t(1) << 31)) return set_error(CCCS_EINVAL, "grid too large");
    CCCS_HIP(hipSetDevice(device));
    DevMem din, dout;
    CCCS_TRY(din.alloc(nin));
    CCCS_TRY(dout.alloc(nout));
    CCCS_HIP(hipMemcpy(din.p, mask, nin, hipMemcpyHostToDevice));
    MaskArgs m;
    m.in = static_cast<const uint8_t*>(din.p);
    m.ix = (int)in_size[0]; m.iy = (int)in_size[1]; m.iz = (int)in_size[2];
    m.is_x = in_start_cm[0]; m.is_y = in_start_cm[1]; m.is_z = in_start_cm[2];
    m.isp_x = in_spacing_cm[0]; m.isp_y = in_spacing_cm[1]; m.isp_z = in_spacing_cm[2];
    m.ox = (int)out_size[0]; m.oy = (int)out_size[1]; m.oz = (int)out_size[2];
    m.os_x = out_start_cm[0]; m.os_y = out_start_cm[1]; m.os_z = out_start_cm[2];
    m.osp_x = out_spacing_cm[0]; m.osp_y = out_spacing_cm[1]; m.osp_z = out_spacing_cm[2];
    m.out = static_cast<uint8_t*>(dout.p);
    dim3 blk(64, 4, 1), grd((m.ox + 63) / 64, (m.oy + 3) / 4, m.oz);
    hipLaunchKernelGGL(mask_resample_kernel, grd, blk, 0, 0, m);
    CCCS_HIP(hipGetLastError());
    CCCS_HIP(hipMemcpy(out, dout.p, nout, hipMemcpyDeviceToHost));
    return CCCS_OK;
}
