/*
 * cccs_oracle.h -- CPU oracle for the CCCS dose path.  TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++17/OpenMP restatement of the reference algorithm (xuqifan897/EndtoEnd, CCCS/):
 * it is the checker the HIP product is compared with, and the CPU baseline that bench.py
 * times.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (endtoend_amd/) never links, includes or calls anything in this directory.
 *
 * Every function states the reference file:line it follows and how it is pinned (closed forms,
 * invariants, reference unit-test known answers); see the test files named per function.
 */
#ifndef CCCS_ORACLE_H_
#define CCCS_ORACLE_H_

#include <stdint.h>
#include "../include/cccs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Counters the oracle records while computing a beam (feed the algorithmic-byte model). */
typedef struct orc_stats {
    uint64_t terma_voxels;      /* calc_bbox voxels                                         */
    uint64_t terma_active;      /* voxels with non-zero fluence (Siddon march executed)     */
    uint64_t siddon_steps;      /* plane crossings processed                                */
    uint64_t rev_samples;       /* REV grid samples over all cones                          */
    uint64_t rev_valid;         /* REV samples with density >= 0                            */
    uint64_t cone_iters;        /* inner line-walk iterations over valid samples             */
    uint64_t cone_iters_all;    /* inner iterations including the walk through invalid ones  */
    uint64_t pvcs_hits;         /* (voxel, cone) pairs that landed inside the REV grid      */
} orc_stats;

/* A1: BEAM::reconfigure (CCCS/DoseCalcIO/beam.cpp:300-331), orientation 0 = ANGLE, 1 = DIRECTION */
int orc_reconfigure(cccs_beam* beam, int orientation);

/* A9: read_spectrum_file (spectra.cpp:7-62), read_kernel (kernel.cpp:27-122),
 * make_poly_kernel (kernel.cpp:123-258), make_cumulative_kernel (kernel.cpp:260-431).
 * poly_total_out: 48*24 poly total weights (may be NULL); cat_out: 5*ntheta*24 cumulative
 * per-category weights (may be NULL). */
int orc_kernel_tables(const char* spectrum_file, const char* kernel_dir, const char* radii_file,
                      int nphi, int ntheta, float* kernel_out, float* radii_out, float* theta_out,
                      float* phi_out, float* spec_out, int* nspec, float* poly_total_out,
                      float* cat_out);
/* make_cumulative_kernel alone, from poly total weights (48 x 24). */
int orc_cumulative_kernel(const float* poly_total, int nphi, int ntheta, float* kernel_out,
                          float* theta_out, float* phi_out);

/* A2: cudaSiddon kernel (DoseCalcAlgo/cudaSiddon.cu:210-308) over the calc_bbox. */
int orc_terma(const cccs_problem* p, const cccs_beam* beam, const float* fmap, int beamletnum,
              float* terma_bbox, orc_stats* stats);
/* A2: cudaRaytrace_device (cudaSiddon.cu:20-115) per end point, with its segment list. */
int orc_siddon_paths(const cccs_problem* p, const float* source, const float* points, int npoints,
                     int max_segments, int32_t* ijk, float* mid, float* len, int32_t* counts,
                     float* rpl);

/* A3 + A4: crop box (dosecalc-beam/nvbbRayConvolve_host.cu:429-499) and findREV (:75-155). */
int orc_beam_geometry(const cccs_problem* p, const cccs_beam* beam, const float* terma_bbox,
                      cccs_beam_geometry* geom);

/* A2-A7: full plain beam (radconvolveTexture, nvbbRayConvolve_host.cu:376-725); dose += . */
int orc_beam_dose(const cccs_problem* p, const cccs_beam* beam, const float* fmap, float* dose,
                  orc_stats* stats);

/* Stage dump for one cone: REV density/terma (A5) and REV cone dose (A6), REV-grid sized
 * (line fastest).  Buffers must hold rev_size product floats (query orc_beam_geometry). */
int orc_beam_cone_stage(const cccs_problem* p, const cccs_beam* beam, const float* fmap, int ray,
                        float* rev_dens, float* rev_terma, float* rev_dose);

/* A7 alone: REV2XYZdose of a caller-supplied REV dose grid of cone `ray` (REV-grid sized, line
 * fastest) into dose (full volume, +=), with the geometry from orc_beam_geometry. */
int orc_rev_to_pvcs(const cccs_problem* p, const cccs_beam* beam, const cccs_beam_geometry* geom, int ray,
                    const float* rev_dose, float* dose);

/* A6 on one line (RowConvolve body, dosecalc-beam/nvbbRayConvolve_device.cu:113-217):
 * dens/terma of length n, kernel row P forward / ntheta-1-P backward; radii = KERN_RADII. */
int orc_row_convolve(const float* dens, const float* terma, int n, const float* kernel, int ntheta,
                     int nradii, const float* radii, int P, float delr, int nphi, float* out);

/* A12 on one line (PackRowConvolve body, dosecalc-beamlet/nvbbRayConvolve_device.cu:284-417). */
int orc_pack_row_convolve(const float* dens, const float* terma, int n, const float* kernel,
                          int ntheta, int nradii, const float* radii, int P, float delr, int nphi,
                          float* out);

/* A10-A15: beamlet path for one beam (dosecalc-beamlet radconvolveTexture, :555-1349), dense
 * columns over the calc_bbox in active-beamlet order.  mode as in cccs.h. */
int orc_beamlet_dense(const cccs_problem* p, const cccs_beam* beam, const float* fmap, int mode,
                      float* out, int max_beamlets, int* nactive, orc_stats* stats);

/* Stage dump of the beamlet path for one cone (oracle instrumentation): batch 0's REV density
 * (cudaBeamletRaytrace, dosecalc-beamlet/nvbbRayConvolve_device.cu:19-242), REV TERMA and REV cone
 * dose (PackRowConvolve, :253-418), line fastest.  rev_size receives the REV grid dims even when
 * cap (floats per buffer) is too small, in which case CCCS_EINVAL is returned. */
int orc_beamlet_cone_stage(const cccs_problem* p, const cccs_beam* beam, const float* fmap, int ray,
                           float* rev_dens, float* rev_terma, float* rev_dose, uint64_t cap,
                           uint32_t* rev_size);

/* Unpack check of the beamlet path (oracle instrumentation): each packed node is filled with
 * coef[0] + (coef[1], coef[2], coef[3]) . X, X the PVCS point the raytrace assigns to the node, and
 * UnpackBEVDosePillar (device.cu:459-521) resamples the grids into dense columns as
 * orc_beamlet_dense does.  An unpack that inverts the raytrace's map returns the linear field at
 * every voxel it writes. */
int orc_beamlet_unpack_linear(const cccs_problem* p, const cccs_beam* beam, const float* fmap,
                              const float* coef, float* out, int max_beamlets, int* nactive);

/* A15: _sparsify (sparse_io.cpp:22-58) of one dense calc_bbox column. Returns nnz, writes at
 * most cap entries. */
/* A10 alone: calcBeamletAnchors (dosecalc-beamlet/nvbbRayConvolve_host.cu:180-280) for fluence
 * index idx: entry/exit of the beamlet axis through the calc_bbox, its isocentre and its
 * divergence angles {theta, phi}.  Returns CCCS_EGEOM when the axis misses the box. */
int orc_beamlet_anchor(const cccs_problem* p, const cccs_beam* beam, int idx, float* start3, float* end3,
                       float* biso3, float* angles2);

/* Preprocess: PTV projection onto fluence maps (findFluenceProjection + cudaRaycast,
 * dosecalc-preprocess/cudaMemFunctions.cu:308-393, kernels.cuh:136-224); maps back to back. */
int orc_fluence_projection(const cccs_problem* p, const uint8_t* mask, const cccs_beam* beams, int nbeams,
                           float* out);

/* A0 (FCBB): fluence-map convolution, zero outside, same size, j outer / i inner summation. */
int orc_fluence_convolve(const float* f, int nmaps, int w, int h, const float* k, int kw, int kh, float* out);

/* ROI reduction of one sparse column (_reduce_by_roi, DoseCalcIO/sparse_io.cpp:62-158).  ROI r:
 * byte mask over its crop box (x fastest).  Returns the reduced count (writes at most cap). */
int64_t orc_reduce_by_roi(const uint64_t* keys, const float* vals, int64_t nnz, const uint32_t* size, int nroi,
                          const uint8_t* const* masks, const uint32_t* crop_start, const uint32_t* crop_size,
                          uint64_t* out_keys, float* out_vals, int64_t cap, uint64_t* row_block_sizes);

int64_t orc_sparsify(const float* dense, const uint32_t* size, const uint32_t* crop_start,
                     const uint32_t* crop_size, float thresh, uint64_t* keys, float* vals,
                     int64_t cap, float* perc_nonzero, float* perc_dropped);

/* DICOM ingestion (dosecalc-preprocess with --dicom).
 * cudaCreateTexIso + cudaMakeIsotropicWithLUT / deviceHU2dens (cudaMemFunctions.cu:83-216,
 * kernels.cuh:13-135): output voxel (X, Y, Z) samples the CT at texel coordinates
 * (X iso / vx, Y iso / vy, Z iso / vz) + 0.5 with the texture unit's linear filter (8-bit
 * fractional weights, zero border), then the lookup table (nlut >= 2) or (HU + 1024) / 1024 >= 0.05. */
int orc_ct_to_density(const float* hu, const uint32_t* size, const float* voxel_cm, float iso, const float* lut_hu,
                      const float* lut_dens, int nlut, float* out, uint32_t* out_size);
/* cudaResample (point filter, zero border) + cudaThreshold(1) (CudaUtilities/array_sum.cuh:13-58). */
int orc_mask_resample(const uint8_t* in, const uint32_t* in_size, const float* in_start, const float* in_sp,
                      const uint32_t* out_size, const float* out_start, const float* out_sp, uint8_t* out);
/* getROIExtents / getROICentroid (DoseCalcIO/roi.cpp:245-345). */
int orc_roi_extents(const float* pts_mm, uint64_t npts, const uint32_t* size, const float* start, const float* sp,
                    uint32_t* crop_start, uint32_t* crop_size);
int orc_roi_centroid(const uint8_t* mask, const uint32_t* size, const float* start, const float* sp, float* c);
/* Per-slice polygon fill of generateContourMask (cudaMemFunctions.cu:218-255): 8-connected edge
 * lines plus scanline interiors (pixels ceil(left)..floor(right) between sorted crossings of the
 * edges spanning [y0, y1)), OR-ed per slice.  Parity unpinned against OpenCV's fillPoly (absent
 * here) beyond the shapes the tests fix. */
int orc_contour_rasterize(const float* pts_mm, const uint32_t* counts, int nsub, const uint32_t* size,
                          const float* start, const float* sp, uint8_t* mask);

#ifdef __cplusplus
}
#endif
#endif
