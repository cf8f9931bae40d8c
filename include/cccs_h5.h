/*
 * cccs_h5.h -- C-ABI of the HDF5 file layer (libcccs_h5.so): the reference's HDF5 outputs and
 * inputs, written and read through the HDF5 C library.  Host code only; no device is touched.
 *
 * Plain C types only.  Every entry point returns CCCS_OK (0) or a negative CCCS_E* status from
 * cccs.h; cccs_h5_last_error() gives the message of the calling thread's last failure.  The
 * HDF5 library is not assumed thread-safe: every call holds one process-wide lock.
 *
 * Reference interfaces replaced (paths relative to the xuqifan897/EndtoEnd tree):
 *   Dose.h5               Volume<float>::writeToFile          CCCS/DoseCalcIO/volume.h:118-150
 *                         _write_file_version                  CCCS/DoseCalcIO/fmaps_io.cpp:168-184
 *   volume input          Volume<T>::readFromFile              CCCS/DoseCalcIO/volume.h:84-116,168-182
 *   Dose_Coefficients.h5  SparsifyManager::init                CCCS/dosecalc-beamlet/sparsify_manager.cpp:176-221
 *                         BaseW2FWorker::write_beam_metadata,  sparsify_manager.cpp:44-97
 *                         W2FWorker_Central::write_beamlet     sparsify_manager.cpp:101-120
 *                         _write_patient_metadata,             CCCS/DoseCalcIO/sparse_io.cpp:178-326
 *                         _write_beam_metadata, _write_beamlet_metadata, _write_beamlet_to_hdf5
 *                         BEAM::_writeToHDF5                   CCCS/DoseCalcIO/beam.cpp:134-298
 *   fluence-map input     read_fmaps                           CCCS/DoseCalcIO/fmaps_io.cpp:10-165
 *   ROI masks             ROIMaskList::writeToFile/readFromFile CCCS/DoseCalcIO/roi.cpp:16-95,107-185
 *                         ArrayProps::_writeToHDF5/_readFromHDF5 CCCS/DoseCalcIO/io_data_structs.h:80-134
 */
#ifndef CCCS_H5_H_
#define CCCS_H5_H_

#include "cccs.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* cccs_h5_last_error(void);

/* ---- volumes ------------------------------------------------------------------------------ */
/* Dose.h5: group /filetype {ftmagic 0x2A, ftversionmajor 1, ftversionminor 0} (U8 attributes)
 * and dataset `dset_name` (NULL = "volume"), IEEE_F32LE, dims [Nz][Ny][Nx] (x fastest, as in
 * memory), with attributes dicom_start_cm[3] and voxel_size_cm[3] (IEEE_F32LE).  The file is
 * truncated. */
int cccs_h5_write_volume(const char* path, const char* dset_name, const float* data, const uint32_t size[3],
                         const float start[3], const float voxel[3]);

/* Reads dataset `dset_name` (NULL = "data", the reference reader's default) of a volume file as
 * float (any numeric file type is converted) with its dicom_start_cm and voxel_size_cm
 * attributes.  size/start/voxel are always filled; `out` (capacity `cap` floats) is filled when
 * non-NULL and large enough, else CCCS_EINVAL.  Pass out = NULL to query the size. */
int cccs_h5_read_volume(const char* path, const char* dset_name, float* out, size_t cap, uint32_t size[3],
                        float start[3], float voxel[3]);

/* ---- beamlet coefficients (Dose_Coefficients.h5) ------------------------------------------ */
/* The per-calculation header (HEADER_PATIENT, sparse_io.h), written once as attributes of
 * /calc_specs. */
typedef struct cccs_h5_calc_specs {
    const char*  version;            /* dosecalc_version string                               */
    int          nbeams;             /* N_beams                                                */
    float        sparsity_thresh;
    const char*  beam_spectrum;      /* spectrum file name                                    */
    const char*  target_structure;
    /* ROI reduction (NULL/0 when the columns are not reduced): roi_order and
     * row_block_capacities (the count of voxels equal to 1 in each ROI mask, roi.cpp:107-113) */
    int          nroi;
    const char* const* roi_names;
    const uint64_t*    roi_capacities;
} cccs_h5_calc_specs;

/* Per-beam strings of the BEAM record (beam.cpp:29-89,248-263); NULL fields default to what an
 * unlabelled BEAM writes: "unspec", "unspec" and "unknown". */
typedef struct cccs_h5_beam_labels {
    const char* isocenter_type;
    const char* isocenter_location;
    const char* orientation_type;
} cccs_h5_beam_labels;

typedef struct cccs_h5_coeff_file cccs_h5_coeff_file;

/* Creates (truncates) the file and writes /filetype {0x2B, 1, 2}, /calc_specs from `specs` and
 * `prob` (dicom_start_cm, full_dicom_size, voxel_size_cm, convlat_cm, convstep_cm,
 * calc_bbox_start, calc_bbox_size, penumbra_cm, nphi, ntheta, nradii, kernel_extent_cm) and the
 * empty groups /beams, /beams/metadata, /beams/data.  compress_lvl < 0 writes contiguous
 * datasets; 0..9 writes deflated chunks of min(N, chunksize) entries (sparse_io.cpp:181-190). */
int cccs_h5_coeff_create(const char* path, const cccs_problem* prob, const cccs_h5_calc_specs* specs,
                         int compress_lvl, uint64_t chunksize, cccs_h5_coeff_file** out);

/* Output layout of the columns (WRITE_STRATEGY, dosecalc-beamlet/sparsify_manager.h:25-28):
 * CENTRAL writes every column's datasets into the file; PER_BEAMLET writes each column to its own
 * HDF5 file "<path>_external/beam_%05d/beam%05d_beamlet%05d.bmlt" (beamlet attributes and the
 * lindex/coeffs datasets at its root) and keeps in the central group the attributes and two external
 * links, lindex and coeffs, to "<basename(path)>_external/beam_.../...bmlt" /lindex and /coeffs
 * (sparsify_manager.cpp:129-174; the sidecar directory is removed and created afresh, :196-204).
 * Call after cccs_h5_coeff_create, before any beam is declared. */
#define CCCS_H5_WRITE_CENTRAL     0
#define CCCS_H5_WRITE_PER_BEAMLET 1
int cccs_h5_coeff_set_write_strategy(cccs_h5_coeff_file* f, int strategy);

/* Declares the beams of the next cccs_beamlet_dose request: request beam i gets uid
 * uid_base + i.  Each beam with at least one beamlet of fluence > 0 gets
 * /beams/metadata/beam_%05d (beam_uid, N_beamlets, beam_specs compound, isocenter_type,
 * isocenter_location, orientation_type, fmap_weights [fmap_x][fmap_y]) and an empty
 * /beams/data/beam_%05d; a uid already in the file is not written again.  labels may be NULL. */
int cccs_h5_coeff_add_beams(cccs_h5_coeff_file* f, const cccs_beam* beams, int nbeams, const float* const* fluence,
                            int uid_base, const cccs_h5_beam_labels* labels);

/* Writes one column as /beams/data/beam_%05d/beamlet_%05d (uid of the column's request beam,
 * beamlet_index): attributes beamlet_uid, N_coeffs, perc_nonzero, perc_dropped (+ row_block_sizes
 * and perc_reducedrop for reduced columns) and datasets lindex (STD_U64LE) and coeffs
 * (IEEE_F32LE).  A column with no entries writes nothing, as the reference skips it. */
int cccs_h5_coeff_add_column(cccs_h5_coeff_file* f, const cccs_sparse_column* column);

/* A cccs_sparse_sink: pass it to cccs_beamlet_dose with `user` = the cccs_h5_coeff_file*; the
 * columns stream into the file as they are produced.  Returns non-zero (aborting the dose call)
 * when a write fails; the message is in cccs_h5_last_error(). */
int cccs_h5_coeff_sink(void* user, const cccs_sparse_column* column);

/* Columns and coefficients written so far. */
int cccs_h5_coeff_counts(const cccs_h5_coeff_file* f, uint64_t* columns, uint64_t* coeffs);

/* Flushes and closes the file and frees `f` (also after a failed write). */
int cccs_h5_coeff_close(cccs_h5_coeff_file* f);

/* ---- fluence-map files (fmaps.h5) ------------------------------------------------------------ */
/* Writes the fluence-map file dosecalc reads with --fmaps (the format of Fmaps.generate,
 * extern/python/fmaps.py:85-123): /filetype {0x2C, 1, 1} and /beams/metadata/beam_%05d for every
 * beam i (uid i; N_beamlets = non-zero weights), with the same beam attributes as the
 * coefficient file.  labels may be NULL ("man", "unspec", "auto", as fmaps.py writes). */
int cccs_h5_write_fmaps(const char* path, const cccs_beam* beams, int nbeams, const float* const* fluence,
                        const cccs_h5_beam_labels* labels);


/* Reads the beams of an fmaps file or a coefficient file (/beams/metadata/beam_* when present,
 * else /beams/beam_*), sorted by beam_uid: each beam's record (reconfigured source/direction as
 * stored) and its fluence map (fmap_weights, fmap_x*fmap_y floats in stored order, which is x
 * fastest).  Two calls: with beams = NULL, *nbeams and *nweights report the counts; then beams
 * (cap_beams) and weights (cap_weights floats, maps back to back in beam order) are filled.
 * uids (may be NULL) receives the beam_uid of each beam. */
int cccs_h5_read_fmaps(const char* path, cccs_beam* beams, int* uids, int cap_beams, int* nbeams, float* weights,
                       size_t cap_weights, size_t* nweights);

/* ---- ROI masks (roi_list.h5) ---------------------------------------------------------------- */
/* One group per ROI named by the ROI, with attributes name (string) and index (1-based U16), a
 * subgroup ArrayProps (size, crop_size, crop_start: 3 x U16 array attributes) and a dataset
 * mask (STD_U8LE, crop_size voxels, x fastest).  size is the CT volume size. */
int cccs_h5_write_roi_list(const char* path, int nroi, const char* const* names, const uint8_t* const* masks,
                           const uint32_t size[3], const uint32_t* crop_start, const uint32_t* crop_size);

/* Reads a ROI list in index order.  Two calls: with masks = NULL, *nroi and *nbytes report the
 * counts; then names (cap_roi entries of 64 chars, NUL-terminated, truncated), crop_start and
 * crop_size (3*cap_roi each) and masks (cap_bytes bytes, masks back to back) are filled.  size
 * receives the volume size of the first ROI. */
int cccs_h5_read_roi_list(const char* path, int cap_roi, int* nroi, char (*names)[64], uint32_t size[3],
                          uint32_t* crop_start, uint32_t* crop_size, uint8_t* masks, size_t cap_bytes, size_t* nbytes);

#ifdef __cplusplus
}
#endif

#endif  /* CCCS_H5_H_ */
