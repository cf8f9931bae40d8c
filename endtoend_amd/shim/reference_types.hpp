// reference_types.hpp -- the reference's host types that its dose entry points take, for the
// source-level drop-in of endtoend_amd/shim/reference_shim.cpp.
//
// Built inside the reference tree (-DCCCS_SHIM_REFERENCE_HEADERS), the reference's own headers
// are used: include/server/brain_defs.h, DoseCalcIO/kernel.h, DoseCalcIO/beam.h,
// DoseCalcIO/io_data_structs.h.  Built here, the same types are re-declared with the same data
// members in the same order (member functions the entry points do not call are left out, which
// does not change a layout: none of these types has virtual members), so the shim compiles and
// is tested without CUDA, HDF5's C++ API or Boost:
//   CONSTANTS, SHM_DATA        include/server/brain_defs.h:14-97
//   KERNEL, MONO_KERNELS       DoseCalcIO/kernel.h:9-12,23-59
//   BEAM                       DoseCalcIO/beam.h:16-95
//   HEADER_PATIENT             DoseCalcIO/io_data_structs.h:46-69
// float2/float3/uint2/uint3 follow CUDA's vector types (8-byte alignment for the two-wide ones).
#pragma once

#ifdef CCCS_SHIM_REFERENCE_HEADERS
#include "server/brain_defs.h"
#include "DoseCalcIO/beam.h"
#include "DoseCalcIO/kernel.h"
#include "DoseCalcIO/io_data_structs.h"
#else
#include <cstdint>
#include <string>
#include <vector>

typedef unsigned int uint;
typedef unsigned short ushort;
struct float3 { float x, y, z; };
struct uint3 { unsigned int x, y, z; };
struct alignas(8) float2 { float x, y; };
struct alignas(8) uint2 { unsigned int x, y; };

#define MAX_KERNELS 30
#define N_KERNEL_RADII 24
#define N_KERNEL_ANGLES 48
#define N_KERNEL_CATEGORIES 5

class BEAM {
  public:
    enum class ORIENT_T { ANGLE, DIRECTION, UNSPEC };
    enum class ISO_T { UNSPEC, PTV_CENTROID, MANUAL };
    enum class ISO_LOC_T { IN_PTV, OUT_PTV, UNSPEC };

    uint uid;
    float azimuth;
    float zenith;
    float coll;
    float3 source;
    float3 direction;
    ORIENT_T orientation_type;
    float3 isocenter;
    ISO_T isocenter_type;
    ISO_LOC_T isocenter_location;
    float sad;
    uint2 fmap_size;
    float2 beamlet_size;
    std::string fmap_fname;
    std::vector<float> fluence_map;
};

struct CONSTANTS {
    unsigned int nphi;
    unsigned int ntheta;
    unsigned int nradii;
    float* conv_theta_deg;
    float* conv_phi_deg;
    uint ss_factor;
    float rev_latspacing;
    float rev_longspacing;
    float kernel_extent;
    float penumbra;
    float beamhard_correct;
    unsigned int beam_count;
    BEAM beam;
    char target_structure[300];
    char beam_spec[300];
    bool reduce = false;
    float3 start;
    float3 voxel;
    uint3 size;
    int nvoxels() const { return size.x * size.y * size.z; }
    uint3 max_rev_size;
    uint3 calc_bbox_start;
    uint3 calc_bbox_size;
    int bbox_nvoxels() const { return calc_bbox_size.x * calc_bbox_size.y * calc_bbox_size.z; }
    int calc_bbox_precomp_2d_size;
};

struct SHM_DATA {
    unsigned int size_data;
    float* density;
    float* kernel_array;
    float* radial_boundary;
};

class KERNEL {
  public:
    std::string kernel_file;
    int nradii;
    int ntheta;
    float radial_boundary[N_KERNEL_RADII + 1];
    float angular_boundary[N_KERNEL_ANGLES + 1];
    std::vector<float> matrix[N_KERNEL_CATEGORIES];
    std::vector<float> total_matrix;
};

struct MONO_KERNELS {
    int nkernels;
    std::string spectrum_file;
    float energy[MAX_KERNELS];
    float fluence[MAX_KERNELS];
    float mu[MAX_KERNELS];
    float mu_en[MAX_KERNELS];
    KERNEL kernel[MAX_KERNELS];
};

struct HEADER_PATIENT {
    ushort N_beams;
    float3 dicom_start_cm;
    uint3 full_dicom_size;
    float3 voxel_size_cm;
    float rev_latspacing_cm;
    float rev_longspacing_cm;
    uint3 bbox_start;
    uint3 bbox_size;
    float penumbra_cm;
    float sparsity_thresh;
    ushort nphi;
    ushort ntheta;
    ushort nradii;
    float kernel_extent_cm;
    std::string beam_spectrum;
    std::string target_structure;
    std::vector<std::string> roi_order{};
    std::vector<uint64_t> row_block_capacities{};
};
#endif
