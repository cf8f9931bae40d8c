// reference_shim.hpp -- the reference's in-process dose entry points, by their own names and
// signatures, on top of the C-ABI of include/cccs.h (and include/cccs_h5.h for the coefficient
// file).  It replaces the host .cu file that the reference's main.cu includes
// (dosecalc-beam/main.cu:29, dosecalc-beamlet/main.cu:40); see INTEGRATION.md for the switch.
//
// One build per executable, as in the reference: libcccs_shim_beam.so (dosecalc-beam) and, with
// CCCS_SHIM_BEAMLET defined, libcccs_shim_beamlet.so (dosecalc-beamlet, which also gets the
// SparsifyManager that main.cu drives).
#pragma once

#include <stdexcept>
#include <string>

#include "reference_types.hpp"

// The reference throws this when a REV grid exceeds max_rev_size
// (dosecalc-beam/nvbbRayConvolve_host.cu:71-74; beamlet :94-102 retries with more batches, which
// the library does by itself).
class ArraySizeError : public virtual std::runtime_error {
  public:
    ArraySizeError(const std::string& str) : std::runtime_error(str) {}
};

// dosecalc-beam/nvbbRayConvolve_host.cu:158-169 (beamlet :283-294): uploads density, kernel
// table, radii, spectrum and cone angles of device `deviceid` (HIP ordinal gpuid).  Returns 1.
int initCudaConstandTex(SHM_DATA* datavols, MONO_KERNELS* mono, CONSTANTS* constants, int nrays, int deviceid,
                        int gpuid, bool verbose = false, bool timing = false, bool debug_write = false);

#ifndef CCCS_SHIM_BEAMLET
// dosecalc-beam/nvbbRayConvolve_host.cu:296-308: adds the dose of every beam (fluence_map of each
// BEAM) into device `deviceid`'s dose volume.  Returns 1; ArraySizeError on an oversized REV grid.
int radconvolveTexture(MONO_KERNELS* mono, CONSTANTS* constants, BEAM* device_beam_arr, int nbeams, int nrays,
                       int deviceid, int gpuid, bool verbose = false, bool timing = false, bool debugwrite = false);
// dosecalc-beam/nvbbRayConvolve_host.cu:745-771
int freeCudaTexture(int ndevices, int* gpuid_arr);
// dosecalc-beam/nvbbRayConvolve_host.cu:774-837: dose = sum over devices of their dose volumes.
int cuda_sum_device_float_dose(float* dose, int ndevices, CONSTANTS* constants, bool timing = false,
                               bool verbose = false);
#else
// dosecalc-beamlet/sparsify_manager.h:25-29
enum class WRITE_STRATEGY { CENTRAL, PER_BEAM, PER_BEAMLET };

struct cccs_h5_coeff_file;

// The part of dosecalc-beamlet's SparsifyManager (sparsify_manager.h:130-190) that main.cu calls
// (dosecalc-beamlet/main.cu:462-463,479-480,519).  The library hands it finished sparse columns
// (sparsified on the device), which it writes straight into the coefficient file: there are no
// dense-column queues or sparsify threads to run.  CENTRAL and PER_BEAM write one file; PER_BEAMLET
// writes each column to its own sidecar file, linked from the central file.
class SparsifyManager {
  public:
    SparsifyManager() {}
    ~SparsifyManager();
    void init(const std::string& h5filename, HEADER_PATIENT& patient_header, unsigned int num_srworkers = 1,
              unsigned int num_w2fworkers = 1, float sparsity_thresh = 1e-4f,
              WRITE_STRATEGY write_strategy = WRITE_STRATEGY::CENTRAL);
    void activate();
    void deactivate();
    int get_num_srworkers() { return (int)m_srworkers; }
    int get_num_w2fworkers() { return 1; }
    WRITE_STRATEGY get_write_strategy() { return m_strategy; }

    // used by radconvolveTexture
    float threshold() const { return m_thresh; }
    cccs_h5_coeff_file* file() const { return m_file; }

  private:
    cccs_h5_coeff_file* m_file = nullptr;
    float m_thresh = 1e-4f;
    unsigned int m_srworkers = 1;
    WRITE_STRATEGY m_strategy = WRITE_STRATEGY::CENTRAL;
    bool m_active = false;
    HEADER_PATIENT m_header;
};

// Defined by dosecalc-beamlet/main.cu:49.
extern SparsifyManager sparsifymanager;

// dosecalc-beamlet/nvbbRayConvolve_host.cu:444-456: the sparse columns of every active beamlet of
// every beam go into sparsifymanager's coefficient file, under each BEAM's uid.  Returns 1.
int radconvolveTexture(MONO_KERNELS* mono, CONSTANTS* constants, BEAM* device_beam_arr, int nbeams, int nrays,
                       int deviceid, int gpuid, int num_unpack_streams, bool verbose = false, bool timing = false,
                       bool debugwrite = false);
// dosecalc-beamlet/nvbbRayConvolve_host.cu:1412-1413
int freeCudaTexture(int ndevices, int* gpuid_arr, bool debugwrite);
#endif
