/*
 * cccs.h -- C-ABI of the MI355X collapsed-cone convolution/superposition (CCCS) dose engine.
 *
 * Plain C types only: pointers, sizes and POD structs.  Every entry point returns an int status
 * (CCCS_OK == 0, negative on failure) and never throws across the boundary.  Host buffers are
 * owned by the caller; device memory is owned by the context unless a call says otherwise.
 *
 * Reference interfaces replaced (paths relative to the xuqifan897/EndtoEnd tree):
 *   problem description   CONSTANTS / SHM_DATA          CCCS/include/server/brain_defs.h:14-90
 *   beam description      BEAM                          CCCS/DoseCalcIO/beam.h:16-95
 *   context set-up        initCudaConstandTex            CCCS/dosecalc-beam/nvbbRayConvolve_host.cu:158-293
 *                                                        CCCS/dosecalc-beamlet/nvbbRayConvolve_host.cu:283-433
 *   full-beam dose        radconvolveTexture (beam)      CCCS/dosecalc-beam/nvbbRayConvolve_host.cu:296-742
 *   beamlet coefficients  radconvolveTexture (beamlet)   CCCS/dosecalc-beamlet/nvbbRayConvolve_host.cu:444-1409
 *   sparse columns        _sparsify                      CCCS/DoseCalcIO/sparse_io.cpp:22-58
 *   tear-down             freeCudaTexture                CCCS/dosecalc-beam/nvbbRayConvolve_host.cu:745-771
 *   dose read-back        cuda_sum_device_float_dose     CCCS/dosecalc-beam/nvbbRayConvolve_host.cu:774-837
 */
#ifndef CCCS_H_
#define CCCS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------ */
#define CCCS_OK          0
#define CCCS_EINVAL     -1  /* bad argument or inconsistent problem description                */
#define CCCS_EGEOM      -2  /* geometry outside limits: REV grid > max_rev_size, line > limit,
                               empty/oversized crop box (the reference throws ArraySizeError /
                               runtime_error here: nvbbRayConvolve_host.cu:148-154, 494-498, 567) */
#define CCCS_ENOMEM     -3  /* device allocation failed or request exceeds the HBM budget      */
#define CCCS_EIO        -4  /* file could not be read or parsed                                */
#define CCCS_EDEVICE    -5  /* HIP runtime error                                               */
#define CCCS_EBUSY      -6  /* serving queue full / request not finished yet                   */

/* ---- problem description (CONSTANTS + SHM_DATA, brain_defs.h:14-90) ---------------------- */
typedef struct cccs_problem {
    uint32_t size[3];          /* PVCS volume dims Nx Ny Nz (x fastest in memory)              */
    float    voxel[3];         /* voxel size [cm]                                               */
    float    start[3];         /* coords of the CENTRE of voxel 0 [cm]                          */
    uint32_t bbox_start[3];    /* calc_bbox start (voxels)                                      */
    uint32_t bbox_size[3];     /* calc_bbox size (voxels)                                       */
    uint32_t max_rev_size[3];  /* REV grid limit; 0 = derive as the beam path does (:248-258)  */
    float    rev_latspacing;   /* convolution lateral ray spacing [cm]                          */
    float    rev_longspacing;  /* convolution step [cm]                                         */
    float    kernel_extent;    /* kernel truncation used for pillar sizing [cm]                 */
    float    penumbra;         /* crop-box margin [cm]                                          */
    uint32_t ss_factor;        /* TERMA supersampling factor (beam path)                        */
    uint32_t nphi;             /* azimuthal cone count                                          */
    uint32_t ntheta;           /* polar cone count (even, divides 48)                           */
    uint32_t nradii;           /* radial samples of the cumulative kernel (24)                  */
    float    beamhard_correct; /* TERMA scale (the reference fixes it to 1)                     */
    /* host tables, copied by cccs_ctx_create */
    const float* density;        /* Nz*Ny*Nx, g/cc                                               */
    const float* kernel;         /* cumulative kernel, ntheta x nradii, r fastest                */
    const float* radii;          /* nradii radial boundaries [cm]                                */
    const float* conv_theta_deg; /* nphi*ntheta/2 cone polar angles                              */
    const float* conv_phi_deg;   /* nphi*ntheta/2 cone azimuth angles                            */
    uint32_t     nspec;          /* spectrum bins (<= 30)                                        */
    const float* spec_fluence;   /* normalised fluence per bin                                   */
    const float* spec_energy;    /* MeV                                                          */
    const float* spec_mu_en;     /* cm^2/g                                                       */
    const float* spec_mu;        /* cm^2/g                                                       */
} cccs_problem;

/* ---- beam description (BEAM, beam.h:16-95) ------------------------------------------------- */
typedef struct cccs_beam {
    float    azimuth;          /* gantry [rad]                                                  */
    float    zenith;           /* couch  [rad]                                                  */
    float    coll;             /* collimator [rad]                                              */
    float    source[3];        /* x-ray source [cm]                                             */
    float    direction[3];     /* unit beam direction                                           */
    float    isocenter[3];     /* [cm]                                                          */
    float    sad;              /* source-axis distance [cm]                                     */
    uint32_t fmap_size[2];     /* fluence map dims (x fastest)                                  */
    float    beamlet_size[2];  /* [cm]                                                          */
} cccs_beam;

/* Per-beam geometry the plain path derives on the host (crop box + REV grids), exposed so
 * tests can compare it with the oracle.  REV arrays hold one entry per cone direction. */
typedef struct cccs_beam_geometry {
    float    lim_min[3];       /* BEV crop-box origin in PVCS [cm]                              */
    uint32_t bev_size[3];      /* BEV crop-box size (index units)                               */
    float    bev_min[3];       /* BEV-index min of non-zero TERMA minus penumbra                */
    float    bev_max[3];       /* BEV-index max of non-zero TERMA plus penumbra                 */
    int      nrays;
    float    rev_min[48][3];   /* REV min coords per cone (<= 48 cone lines)                    */
    uint32_t rev_size[48][3];  /* REV grid dims per cone (line length first)                    */
} cccs_beam_geometry;

/* Sparse beamlet column: full-volume linear keys + FP32 values, in the reference's emission
 * order (i outer, j, k inner; sparse_io.cpp:34-53). */
typedef struct cccs_sparse_column {
    int       beam_index;      /* position of the beam in the request                          */
    int       beamlet_index;   /* linear fluence-map index of the beamlet                      */
    uint64_t  nnz;
    const uint64_t* keys;      /* valid only during the sink callback                          */
    const float*    values;
    float     perc_nonzero;
    float     perc_dropped;
    /* ROI-reduced columns (cccs_set_roi_masks): one block per ROI, row_block_sizes[r] entries in
     * block r; nblocks = 0 and row_block_sizes = NULL otherwise.  perc_reducedrop is the
     * percentage of calc_bbox voxels the reduction removed (sparse_io.cpp:154). */
    int       nblocks;
    const uint64_t* row_block_sizes;
    float     perc_reducedrop;
} cccs_sparse_column;

/* Called once per beamlet column; return 0 to continue, non-zero to abort the call. */
typedef int (*cccs_sparse_sink)(void* user, const cccs_sparse_column* column);

typedef struct cccs_ctx cccs_ctx;

/* Selects the beamlet pipeline.
 *   CCCS_BEAMLET_FAITHFUL : batched pipeline; every footprint REV sample of every cone gets its
 *                           own Siddon TERMA, as the reference computes it
 *                           (dosecalc-beamlet/nvbbRayConvolve_device.cu:170-211).
 *   CCCS_BEAMLET_PLAIN    : the reference dataflow (one batch at a time, per-cone launches over
 *                           whole REV grids, dense per-beamlet unpack copied to the host and
 *                           sparsified there); the baseline of the batched pipeline.
 * Any other value is rejected with CCCS_EINVAL. */
#define CCCS_BEAMLET_FAITHFUL 0
#define CCCS_BEAMLET_PLAIN    2

/* ---- context: replaces initCudaConstandTex / freeCudaTexture --------------------------------
 * Uploads density, kernel table, radii, spectrum and cone angles to the device and keeps them
 * resident until cccs_ctx_destroy.  `device` is the HIP ordinal. */
int cccs_ctx_create(const cccs_problem* problem, int device, cccs_ctx** out);
int cccs_ctx_destroy(cccs_ctx* ctx);
/* HBM bytes held by the context (resident state plus cached work buffers). */
int cccs_ctx_memory(const cccs_ctx* ctx, uint64_t* resident_bytes, uint64_t* scratch_bytes);
/* Launch-schedule knob (no counterpart in the reference; results do not depend on it beyond FP32
 * rounding).  Keys:
 *   "cone_ramp_min"  full-beam cone lines with at least this many valid samples use the
 *                    segment-sum walk, O(nradii) per sample, instead of the sequential walk,
 *                    O(line) per sample (default 64; a value above the longest line disables it).
 *   "cone_class_len0".."cone_class_len2"
 *                    upper stored-line lengths of the first three cone-kernel launch classes
 *                    (defaults 96, 160, 320; the last class takes the longer lines).  Each class
 *                    launch sizes its LDS for its longest line.
 *   "bl_streams"     cone directions of the beamlet path in flight at once, 1 to 4 (default 4:
 *                    consecutive directions' sample, Siddon and walk overlap on their own
 *                    streams; the gathers stay in direction order).
 *   "bl_sample_waves", "bl_walk_waves"
 *                    waves of the beamlet path's persistent per-line kernels (defaults 262144 and
 *                    32768; each wave takes a stride of lines).
 *   "bl_walk_split"  the beamlet walk runs one launch per line-length class (up to this many
 *                    stored samples, up to twice that, ..., the rest), each with its LDS sized
 *                    for its own longest line (default 256; 0 = one launch).
 *   "cone_class_threads0".."cone_class_threads3"
 *                    threads per cone line in each class: 64, 128 or 256 (defaults 64, 64, 128,
 *                    128). */
int cccs_ctx_set_schedule(cccs_ctx* ctx, const char* key, int value);
int cccs_ctx_get_schedule(const cccs_ctx* ctx, const char* key, int* value);

/* Autotuner (no counterpart in the reference): times the given workload under candidate values of
 * the schedule knobs, one knob at a time over two passes (coordinate descent), and leaves the
 * fastest setting on the context.  CCCS_TUNE_BEAM tunes the full-beam path's cone-kernel knobs on
 * device-resident beam doses of `beams`; CCCS_TUNE_BEAMLET tunes the beamlet path's wave counts and
 * cone directions in flight (bl_streams) on device-resident faithful columns.  Results do not depend on the knobs beyond FP32 rounding. */
#define CCCS_TUNE_BEAM    0
#define CCCS_TUNE_BEAMLET 1
typedef struct cccs_tune_report {
    int    nkeys;
    char   key[16][32];
    int    value[16];          /* chosen values                                                */
    double default_ms;         /* one call of the workload with the knobs found on entry        */
    double best_ms;            /* one call with the chosen knobs                                */
    int    trials;             /* timed calls made                                              */
} cccs_tune_report;
int cccs_autotune(cccs_ctx* ctx, int kind, const cccs_beam* beams, int nbeams, const float* const* fluence,
                  cccs_tune_report* report);

/* ---- full-beam dose: replaces radconvolveTexture (dosecalc-beam) + cuda_sum_device_float_dose.
 * Computes the dose of `nbeams` beams and adds their sum into `dose` (Nz*Ny*Nx, host memory).
 * fluence[b] points at beam b's fluence map (fmap_size[1] rows of fmap_size[0]). */
int cccs_beam_dose(cccs_ctx* ctx, const cccs_beam* beams, int nbeams,
                   const float* const* fluence, float* dose);

/* Device-resident variant: d_fluence holds the maps back to back (beam order), d_dose is a
 * device buffer of Nz*Ny*Nx floats that is accumulated into.  Runs on the context stream;
 * returns after the work is enqueued and the host-side geometry is resolved. */
int cccs_beam_dose_device(cccs_ctx* ctx, const cccs_beam* beams, int nbeams,
                          const float* d_fluence, float* d_dose);

/* No-allocation variant of cccs_beam_dose_device (no counterpart in the reference; for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and result; it calls no hipMalloc/hipFree/hipHostMalloc.  Its
 * device scratch is a fixed 256 MiB array per GPU, held by the library until the process ends; a
 * beam whose scratch (TERMA, density/TERMA pairs, line table, compact REV store) exceeds it fails
 * with CCCS_ENOMEM before any dose of that beam is added.  Calls on one GPU share that array, so
 * calls that overlap on one GPU (any contexts, any threads) race; per-kernel profiling must be off
 * (CCCS_EINVAL otherwise).  The context itself comes from cccs_ctx_create beforehand. */
int cccs_beam_dose_device_noalloc(cccs_ctx* ctx, const cccs_beam* beams, int nbeams,
                                  const float* d_fluence, float* d_dose);

/* ---- angle cache (no counterpart in the reference; for serving loops and optimisers that ask
 * for beams from a fixed candidate set) ---------------------------------------------------------
 * The unit-fluence TERMA of a full-beam dose depends on the beam's source and SAD and on the
 * problem, not on its fluence map or collimator.  A cache holds, per distinct (source, SAD) of
 * `beams`, that TERMA over the calc_bbox in device memory (4 * bbox voxels bytes each; entries are
 * added in list order while they fit in max_bytes).  A context using a cache takes a request's
 * TERMA as the voxel's fluence times the cached entry, bit-identical to the march it replaces;
 * beams with another source, and problems with ss_factor > 1, run the march.  The cache is
 * read-only after creation and may be shared by contexts of the same problem, density and device
 * (cccs_ctx_use_angle_cache rejects another); destroy it only after every context using it has
 * finished its work or dropped it (use NULL). */
typedef struct cccs_angle_cache cccs_angle_cache;
int cccs_angle_cache_create(cccs_ctx* ctx, const cccs_beam* beams, int nbeams, uint64_t max_bytes,
                            cccs_angle_cache** out);
int cccs_angle_cache_info(const cccs_angle_cache* cache, int* nangles, uint64_t* bytes);
int cccs_ctx_use_angle_cache(cccs_ctx* ctx, const cccs_angle_cache* cache);
int cccs_angle_cache_destroy(cccs_angle_cache* cache);

/* Stage outputs of the beam path, for stage-wise parity checks. */
int cccs_beam_terma(cccs_ctx* ctx, const cccs_beam* beam, const float* fluence, int beamletnum,
                    float* terma_bbox);
int cccs_beam_geometry_of(cccs_ctx* ctx, const cccs_beam* beam, const float* fluence,
                          cccs_beam_geometry* geom);

/* Siddon voxel-index lists (bit-exact artefact).  For each of `npoints` end points traces the
 * ray from `source`, writing up to `max_segments` segments per point: ijk[3] voxel index of the
 * segment midpoint, mid[3] the midpoint fetch coordinate, len the segment length [cm].
 * counts[p] receives the segment count (may exceed max_segments; extra segments are dropped)
 * and rpl[p] the radiological path length. */
int cccs_siddon_paths(cccs_ctx* ctx, const float* source, const float* points, int npoints,
                      int max_segments, int32_t* ijk, float* mid, float* len, int32_t* counts,
                      float* rpl);

/* ---- beamlet coefficients: replaces radconvolveTexture (dosecalc-beamlet) + SparsifyManager.
 * Every active beamlet (fluence > 0) of every beam yields one sparse column, passed to `sink`
 * in beam order then active-beamlet order.  threshold: keep v > threshold * max(column).
 * sink may be NULL (not in the plain mode): the columns are then built on the device and not
 * copied out -- the device-resident throughput of the pipeline. 
 * mode: CCCS_BEAMLET_FAITHFUL or CCCS_BEAMLET_PLAIN. */
int cccs_beamlet_dose(cccs_ctx* ctx, const cccs_beam* beams, int nbeams,
                      const float* const* fluence, float threshold, int mode,
                      cccs_sparse_sink sink, void* user);

/* Counting sink for callers that need totals only (throughput runs, dry runs): pass a
 * cccs_sink_totals as `user`; no column data is kept.  No counterpart in the reference. */
/* ROI reduction of beamlet columns (the `reduce` option of dosecalc-beamlet; _reduce_by_roi,
 * DoseCalcIO/sparse_io.cpp:62-158).  ROI r is a byte mask over its crop box: crop_size[3r..3r+2]
 * voxels starting at crop_start[3r..3r+2] on the CT grid, x fastest, non-zero = inside.  Once set,
 * cccs_beamlet_dose delivers every column of the batched modes reduced: block r holds the
 * column's kept entries inside ROI r, in the column's order, keyed (voxels of all ROIs before r) +
 * (rank of the voxel among ROI r's voxels in x-outer, y, z-inner order).  A voxel in two ROIs
 * appears in both blocks.  nroi = 0 turns the reduction off.  The plain mode rejects ROIs. */
int cccs_set_roi_masks(cccs_ctx* ctx, int nroi, const uint8_t* const* masks, const uint32_t* crop_start,
                       const uint32_t* crop_size);

typedef struct cccs_sink_totals {
    uint64_t columns;
    uint64_t nnz;
    double   value_sum;
} cccs_sink_totals;
int cccs_sink_count(void* user, const cccs_sparse_column* column);

/* Dense variant for parity tests: writes the bbox-sized dense column of each active beamlet of
 * ONE beam into out (nactive x bbox voxels, active-beamlet order). */
int cccs_beamlet_dense(cccs_ctx* ctx, const cccs_beam* beam, const float* fluence, int mode,
                       float* out, int max_beamlets, int* nactive);

/* Device bytes a cccs_beamlet_dose call for this one beam would hold while it runs (packed
 * grids, one cone direction's REV store, half steps and line table, sparse-output bound); host
 * planning only, no device work.  No counterpart in the reference. */
int cccs_beamlet_footprint(const cccs_ctx* ctx, const cccs_beam* beam, const float* fluence, uint64_t* bytes);

/* ---- serving: a request loop over device-resident contexts ----------------------------------
 * For the calls a beam-orientation optimiser issues while annealing (single-beam dose, or one
 * beam's beamlet columns).  `slots` contexts (own stream each) run concurrently; a free slot
 * takes every queued beamlet request that fits and runs them as one multi-beam job (continuous
 * batching: requests that arrive while a batch runs join the next one), and full-beam requests
 * one at a time (each fills the device alone).  Queued full-beam requests are taken first, and
 * with two or more slots at most slots-1 run beamlet batches at once, so a full-beam request
 * never waits behind a long beamlet batch.  A request is admitted only while the bytes
 * in flight plus its estimate stay within hbm_budget (a request above the whole budget runs
 * alone).  Results are the same as the direct calls (cccs_beam_dose / cccs_beamlet_dose) made
 * alone.  No counterpart in the reference (its dose calculation is one batch program). */
#define CCCS_REQ_BEAM_DOSE 0
#define CCCS_REQ_BEAMLET   1
typedef struct cccs_server cccs_server;
typedef struct cccs_server_info {
    uint64_t completed;
    double   p50_ms, p99_ms, max_ms;   /* submit-to-completion latency                          */
    uint64_t batches;
    double   mean_batch;               /* requests per batch                                    */
    uint64_t queued;
} cccs_server_info;
int cccs_server_create(const cccs_problem* problem, int device, int slots, uint64_t hbm_budget,
                       int max_batch, cccs_server** out);
/* Copies the beam and its fluence map.  threshold/mode are used by beamlet requests. */
int cccs_server_submit(cccs_server* server, int kind, const cccs_beam* beam, const float* fluence,
                       float threshold, int mode, uint64_t* request_id);
/* Blocks until the request is done, then delivers it: a beam dose into `dose` (full volume, may
 * be NULL), beamlet columns through `sink` (may be NULL).  The request is forgotten afterwards. */
int cccs_server_wait(cccs_server* server, uint64_t request_id, float* dose, cccs_sparse_sink sink,
                     void* user, double* latency_ms, int* batch_size);
int cccs_server_stats(cccs_server* server, cccs_server_info* out);
/* Builds an angle cache of `beams` (see cccs_angle_cache_create) shared by every slot, replacing
 * any earlier one (nbeams 0 drops it).  Only while no request is queued or running (CCCS_EBUSY
 * otherwise); submissions wait until it returns.  The cache is device memory outside hbm_budget,
 * resident until the next preload or cccs_server_destroy.  No counterpart in the reference. */
int cccs_server_preload_angles(cccs_server* server, const cccs_beam* beams, int nbeams, uint64_t max_bytes,
                               int* nangles);
/* Finishes every queued request, then releases the slots. */
int cccs_server_destroy(cccs_server* server);

/* ---- preprocess inputs -------------------------------------------------------------------------
 * Projection of a structure mask onto each beam's fluence map (findFluenceProjection + cudaRaycast,
 * dosecalc-preprocess/cudaMemFunctions.cu:308-393, kernels.cuh:136-224).  `mask` holds
 * size[2]*size[1]*size[0] bytes on the context's CT grid (non-zero = inside the structure).  For
 * each beamlet, nine sub-rays (the cell centre and the points half a beamlet away along each
 * fluence axis and diagonal) run from the source to three times the source-to-cell distance
 * through the trilinearly sampled mask with the TERMA Siddon march; the beamlet is 1 when any
 * sub-ray accumulates a positive path integral, else 0.  `fluence` receives the maps of all beams
 * back to back (fmap_size[1] rows of fmap_size[0] each). */
int cccs_fluence_projection(cccs_ctx* ctx, const uint8_t* mask, const cccs_beam* beams, int nbeams,
                            float* fluence);
/* Fluence-map convolution, the lateral step of fluence-convolution broad-beam (FCBB) dose named by
 * the reference's README (README.md:15-20; no source in the reference): for each of `nmaps` maps
 * of height rows x width, out[y][x] = sum_j sum_i kernel[j][i] * map[y - (j - kh/2)][x - (i - kw/2)]
 * (kernel rows j of kw entries, odd kw, kh <= 31), zero outside the map, the same size out. */
int cccs_fluence_convolve(cccs_ctx* ctx, const float* fluence, int nmaps, int width, int height,
                          const float* kernel, int kw, int kh, float* out);
/* Aperture-ready fill of a fluence map, row by row (fmap_post_apertureready,
 * dosecalc-preprocess/fmapProcessing.cpp:7-46, with its segment bookkeeping: a run of non-zero
 * beamlets is recorded by its last beamlet, so the fill between two runs covers the gap and all
 * but the last beamlet of the next run).  fill_val is 1 in the reference. */
int cccs_fmap_aperture_fill(float* fluence, int width, int height, float fill_val);
/* 1 when every row has at most one recorded run (fmap_is_apertureready, fmapProcessing.cpp:26-32). */
int cccs_fmap_is_aperture_ready(const float* fluence, int width, int height, int* ready);

/* ---- DICOM ingestion (dosecalc-preprocess with --dicom, main.cu:139-320) -----------------------
 * Host readers replace RTImage / RTStruct (CCCS/RTClasses/rtimages.cpp:20-400, rtstruct.cpp:100-380,
 * DCMTK there); files are DICOM Part 10, explicit or implicit VR little endian.
 * cccs_dicom_ct_info/_read: the files of `dir` whose names contain ".dcm" and whose SOP class is CT,
 * MR or RT image, sorted from foot to head by ImagePositionPatient z.  size = {Columns, Rows,
 * slices}; voxel_mm = {PixelSpacing[0], PixelSpacing[0], z1 - z0 (SliceThickness for one slice)}
 * as the reference reads them; start_mm = ImagePositionPatient of the first slice.  The volume is
 * [slice][row][column] in FP32: pixel (two's complement when PixelRepresentation is 1) times
 * RescaleSlope plus RescaleIntercept for CT (16-bit pixels only). */
typedef struct cccs_ct_info {
    uint32_t size[3];
    float    voxel_mm[3];
    float    start_mm[3];
    float    orientation[6];     /* ImageOrientationPatient of the first slice                   */
    float    rescale_slope, rescale_intercept;
    int      sop_type;           /* 1 CT, 2 MR, 3 RT image                                       */
} cccs_ct_info;
int cccs_dicom_ct_info(const char* dir, cccs_ct_info* out);
int cccs_dicom_ct_read(const char* dir, float* hu, uint64_t cap);

/* The first RT Structure Set file of `dir` (any name).  ROI r is item r of the Structure Set ROI
 * Sequence (name, number) with the contours of item r of the ROI Contour Sequence: per contour
 * NumberOfContourPoints x {x, y, z} of ContourData in mm, stored as FP32. */
typedef struct cccs_rtstruct cccs_rtstruct;
int cccs_dicom_rtstruct_open(const char* dir, cccs_rtstruct** out);
int cccs_rtstruct_count(const cccs_rtstruct* s, int* nroi);
int cccs_rtstruct_roi_info(const cccs_rtstruct* s, int roi, char* name, int name_cap, int* roi_number, int* nsub,
                           uint64_t* npoints);
int cccs_rtstruct_roi_points(const cccs_rtstruct* s, int roi, uint32_t* sub_counts, float* points_mm);
/* getROIIndex (DoseCalcIO/roi.cpp:201-215): first ROI whose name equals (exact) or contains
 * `search`; *roi = -1 when none does. */
int cccs_rtstruct_find(const cccs_rtstruct* s, const char* search, int exact, int* roi);
int cccs_rtstruct_close(cccs_rtstruct* s);

/* getROIExtents (DoseCalcIO/roi.cpp:245-316): index box of the contour points on a frame (points in
 * mm, frame in cm), floor/ceil of the extremes, clamped to the frame. */
int cccs_roi_extents(const float* points_mm, uint64_t npoints, const uint32_t frame_size[3],
                     const float frame_start_cm[3], const float frame_spacing_cm[3], uint32_t crop_start[3],
                     uint32_t crop_size[3]);
/* getROICentroid (DoseCalcIO/roi.cpp:319-345): mean of the voxel centres of a mask [cm]. */
int cccs_roi_centroid(const uint8_t* mask, const uint32_t size[3], const float start_cm[3], const float spacing_cm[3],
                      float centroid_cm[3]);
/* The per-slice part of generateContourMask (cudaMemFunctions.cu:218-255) on the CT grid: each
 * contour goes to slice round((z0 - start) / spacing) of its first point, its vertices to
 * floor((x - start) / spacing), and its filled polygon is OR-ed into the slice (edges drawn as
 * 8-connected lines, interior by scanlines between edge crossings). */
int cccs_contour_rasterize(const float* points_mm, const uint32_t* sub_counts, int nsub, const uint32_t ct_size[3],
                           const float ct_start_cm[3], const float ct_spacing_cm[3], uint8_t* mask);
/* load_lookup_table (DoseCalcIO/ctlut.cpp:39-90): "name CT# density [rel]" lines, sorted by CT#.
 * With hunits/massdens NULL only *n is set. */
int cccs_ct_lut_load(const char* path, float* hunits, float* massdens, int cap, int* n);
/* CT volume (FP32 HU, [z][y][x], voxel in cm) resampled on device `device` onto the isotropic grid
 * of iso_voxel_cm (out_size = ceil(voxel * (size / iso)), origin at the CT origin) with linear
 * interpolation (texture-unit weights: 8 fractional bits; zero outside the CT) and converted to
 * density by the sorted lookup table (clamped below its first point, extrapolated above its last),
 * or, with nlut = 0, by (HU + 1024) / 1024 clamped at 0.05 (cudaCreateTexIso and its kernels,
 * cudaMemFunctions.cu:83-216, kernels.cuh:13-135).  density = NULL only reports out_size. */
int cccs_ct_to_density(int device, const float* hu, const uint32_t size[3], const float voxel_cm[3],
                       float iso_voxel_cm, const float* lut_hu, const float* lut_dens, int nlut,
                       float* density, uint64_t cap, uint32_t out_size[3]);
/* Mask on one grid point-sampled onto another (voxel index floor(u + 0.5) of the other grid's
 * position, zero outside) and thresholded at 1 (cudaResample + cudaThreshold,
 * CudaUtilities/array_sum.cuh:13-58), on device `device`. */
int cccs_mask_resample(int device, const uint8_t* mask, const uint32_t in_size[3], const float in_start_cm[3],
                       const float in_spacing_cm[3], const uint32_t out_size[3], const float out_start_cm[3],
                       const float out_spacing_cm[3], uint8_t* out);

/* ---- the reference's file formats (host only) -----------------------------------------------
 * Beam list "az zen [coll] [iso: x y z] [sad: s] [dir: x y z]" in degrees/cm, '#' comments,
 * default SAD 100 cm; each beam is reconfigured (from angles, or from dir: when given).
 * requested <= 0 reads every line.  fmap_size/beamlet_size are left 0 for the caller.
 * Replaces load_beam_list (DoseCalcIO/beam.cpp:369-515). */
int cccs_load_beam_list(const char* path, int requested, cccs_beam* beams, int cap, int* count);
/* omni_beam_list.txt written by the preprocessor; fmap_names receives beam_count strings of
 * name_len bytes (may be NULL).  Replaces load_omni_beam_list (DoseCalcIO/beam.cpp:516-576). */
int cccs_load_omni_beam_list(const char* path, int beam_count, cccs_beam* beams, char* fmap_names,
                             int name_len, int* count);
/* omni-header.txt: fills the scalar fields of `p` (tables and pointers untouched) and clamps the
 * calc_bbox into the volume.  beam_spec/target_structure need 100 bytes each (may be NULL).
 * Replaces load_omni_header (DoseCalcIO/calc_header.cpp:81-161); the writer restates
 * write_omni_header (calc_header.cpp:12-80). */
int cccs_load_omni_header(const char* path, cccs_problem* p, int* beam_count, char* beam_spec,
                          char* target_structure, int* reduce);
int cccs_write_omni_header(const char* path, const cccs_problem* p, int beam_count,
                           const char* beam_spec, const char* target_structure, int reduce);
/* Raw little-endian float32 files: density.raw, fluence maps (fluence.cpp:19-44, exact size
 * required), dose output (<out>.raw, DoseCalcIO/binary_io.h:47-86). */
/* omni_beam_list.txt writer (write_omni_beam_list, DoseCalcIO/beam.cpp:577-607).  The string
 * arrays may be NULL (defaults "unspec", "unspec", "auto" and "fmap-%06d.raw"). */
int cccs_write_omni_beam_list(const char* path, const cccs_beam* beams, int beam_count,
                              const char* const* iso_type, const char* const* iso_loc,
                              const char* const* orient_type, const char* const* fmap_names);
int cccs_read_f32(const char* path, float* out, uint64_t n);
int cccs_write_f32(const char* path, const float* data, uint64_t n);

/* ---- host-side helpers restating the reference's input preparation ------------------------
 * Beam angles -> source/direction (BEAM::reconfigure, beam.cpp:300-331). orientation 0 = from
 * angles, 1 = from direction. */
int cccs_beam_reconfigure(cccs_beam* beam, int orientation);

/* Spectrum + mono-energetic kernels -> cumulative poly kernel and cone angles
 * (read_spectrum_file spectra.cpp:7-62, read_kernel kernel.cpp:27-122, make_poly_kernel
 * kernel.cpp:123-258, make_cumulative_kernel kernel.cpp:260-431).
 * kernel_out: ntheta*nradii; theta_out/phi_out: nphi*ntheta/2; spec_out: 4*nspec rows
 * (fluence, energy, mu_en, mu); radii_out: nradii. */
int cccs_build_kernel_from_files(const char* spectrum_file, const char* kernel_dir,
                                 const char* radii_file, int nphi, int ntheta,
                                 float* kernel_out, float* radii_out, float* theta_out,
                                 float* phi_out, float* spec_out, int* nspec);

/* First half of the above: the poly-energetic kernel (48 polar bins x 24 radii total weights,
 * r fastest), the radial boundaries and the normalised spectrum (4*nspec rows as above). */
int cccs_poly_kernel_from_files(const char* spectrum_file, const char* kernel_dir,
                                const char* radii_file, float* poly_total_out, float* radii_out,
                                float* spec_out, int* nspec);

/* Same, from an already assembled poly-energetic kernel (48 x 24 total weights). */
int cccs_cumulative_kernel(const float* poly_total, int nphi, int ntheta,
                           float* kernel_out, float* theta_out, float* phi_out);

/* ---- device memory helpers (plain pointers on the context's device) ------------------------ */
int cccs_dev_alloc(cccs_ctx* ctx, size_t bytes, void** ptr);
int cccs_dev_free(cccs_ctx* ctx, void* ptr);
int cccs_dev_upload(cccs_ctx* ctx, void* dst, const void* src, size_t bytes);
int cccs_dev_download(cccs_ctx* ctx, void* dst, const void* src, size_t bytes);
int cccs_dev_zero(cccs_ctx* ctx, void* ptr, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (bytes must be a multiple of 16); also the
 * calibration kernel for HBM byte counters. */
int cccs_dev_copy(cccs_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Waits for all work the context has enqueued. */
int cccs_ctx_sync(cccs_ctx* ctx);

/* ---- per-kernel profiling --------------------------------------------------------------------
 * When enabled, every kernel launch of the context is bracketed by HIP events on the stream it
 * runs on, and its algorithmic HBM bytes / FLOPs (from the launch's geometry and the kernel's own
 * work counters) are recorded.  cccs_ctx_profile_read synchronises and returns one row per
 * kernel name, summed over launches since the last reset. */
typedef struct cccs_kernel_stats {
    char     name[32];
    uint64_t launches;
    double   ms_total;         /* sum of event-measured launch durations                         */
    double   alg_bytes;        /* algorithmic HBM bytes (sum over launches)                      */
    double   alg_flops;        /* algorithmic FP32 operations (sum over launches)                */
    uint64_t work_items;       /* kernel-specific count: voxels, samples or walk iterations      */
} cccs_kernel_stats;
int cccs_ctx_profile(cccs_ctx* ctx, int enable);
int cccs_ctx_profile_reset(cccs_ctx* ctx);
int cccs_ctx_profile_read(cccs_ctx* ctx, cccs_kernel_stats* out, int max_rows, int* nrows);

/* Last error message of the calling thread ("" when none). */
const char* cccs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CCCS_H_ */
