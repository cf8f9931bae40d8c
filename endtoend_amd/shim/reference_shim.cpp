// reference_shim.cpp -- the reference's in-process dose entry points (dosecalc-beam and
// dosecalc-beamlet nvbbRayConvolve_host.cu) over the C-ABI of libcccs (include/cccs.h).  The
// declarations and the contract of each function are in reference_shim.hpp.
//
// State per device index, as the reference keeps it in static per-device arrays
// (dosecalc-beam/nvbbRayConvolve_host.cu:29-54): the context (resident density, kernel tables,
// spectrum and cone angles) and, for dosecalc-beam, the device dose accumulator.  Library status
// codes become the reference's error behaviour: CCCS_EGEOM throws ArraySizeError, anything else
// std::runtime_error with the library's message.
#include "reference_shim.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

#include "cccs.h"
#ifdef CCCS_SHIM_BEAMLET
#include "cccs_h5.h"
#endif

namespace {

constexpr int kMaxDevices = 16;
cccs_ctx* g_ctx[kMaxDevices] = {};
#ifndef CCCS_SHIM_BEAMLET
float* g_dose[kMaxDevices] = {};  // device memory of the device's context
#endif

void check(int rc, const char* where) {
    if (rc == CCCS_OK) return;
    const std::string msg = std::string(where) + ": " + cccs_last_error();
    if (rc == CCCS_EGEOM) throw ArraySizeError(msg);
    throw std::runtime_error(msg);
}

int device_slot(int deviceid) {
    if (deviceid < 0 || deviceid >= kMaxDevices) throw std::invalid_argument("device index beyond 16");
    return deviceid;
}

cccs_ctx* ctx_of(int deviceid) {
    cccs_ctx* c = g_ctx[device_slot(deviceid)];
    if (!c) throw std::logic_error("initCudaConstandTex was not called for this device");
    return c;
}

// CONSTANTS + SHM_DATA + MONO_KERNELS -> cccs_problem (brain_defs.h:14-97, kernel.h:51-59).  The
// tables stay owned by the caller; cccs_ctx_create copies them.
cccs_problem problem_of(const SHM_DATA* d, const MONO_KERNELS* mono, const CONSTANTS* c) {
    cccs_problem p;
    std::memset(&p, 0, sizeof(p));
    p.size[0] = c->size.x; p.size[1] = c->size.y; p.size[2] = c->size.z;
    p.voxel[0] = c->voxel.x; p.voxel[1] = c->voxel.y; p.voxel[2] = c->voxel.z;
    p.start[0] = c->start.x; p.start[1] = c->start.y; p.start[2] = c->start.z;
    p.bbox_start[0] = c->calc_bbox_start.x; p.bbox_start[1] = c->calc_bbox_start.y; p.bbox_start[2] = c->calc_bbox_start.z;
    p.bbox_size[0] = c->calc_bbox_size.x; p.bbox_size[1] = c->calc_bbox_size.y; p.bbox_size[2] = c->calc_bbox_size.z;
    p.max_rev_size[0] = c->max_rev_size.x; p.max_rev_size[1] = c->max_rev_size.y; p.max_rev_size[2] = c->max_rev_size.z;
    p.rev_latspacing = c->rev_latspacing;
    p.rev_longspacing = c->rev_longspacing;
    p.kernel_extent = c->kernel_extent;
    p.penumbra = c->penumbra;
    p.ss_factor = c->ss_factor;
    p.nphi = c->nphi;
    p.ntheta = c->ntheta;
    p.nradii = c->nradii;
    p.beamhard_correct = c->beamhard_correct;
    p.density = d->density;
    p.kernel = d->kernel_array;       // cumulative kernel, ntheta rows of nradii
    p.radii = d->radial_boundary;
    p.conv_theta_deg = c->conv_theta_deg;
    p.conv_phi_deg = c->conv_phi_deg;
    p.nspec = (uint32_t)mono->nkernels;
    p.spec_fluence = mono->fluence;
    p.spec_energy = mono->energy;
    p.spec_mu_en = mono->mu_en;
    p.spec_mu = mono->mu;
    return p;
}

cccs_beam beam_of(const BEAM& b) {
    cccs_beam o;
    std::memset(&o, 0, sizeof(o));
    o.azimuth = b.azimuth; o.zenith = b.zenith; o.coll = b.coll;
    o.source[0] = b.source.x; o.source[1] = b.source.y; o.source[2] = b.source.z;
    o.direction[0] = b.direction.x; o.direction[1] = b.direction.y; o.direction[2] = b.direction.z;
    o.isocenter[0] = b.isocenter.x; o.isocenter[1] = b.isocenter.y; o.isocenter[2] = b.isocenter.z;
    o.sad = b.sad;
    o.fmap_size[0] = b.fmap_size.x; o.fmap_size[1] = b.fmap_size.y;
    o.beamlet_size[0] = b.beamlet_size.x; o.beamlet_size[1] = b.beamlet_size.y;
    return o;
}

// Each BEAM carries its own fluence map (beam.h:52); it must hold fmap_size.x * fmap_size.y weights.
void beams_of(const BEAM* arr, int n, std::vector<cccs_beam>& beams, std::vector<const float*>& fl) {
    beams.resize((size_t)n);
    fl.resize((size_t)n);
    for (int i = 0; i < n; i++) {
        beams[(size_t)i] = beam_of(arr[i]);
        if (arr[i].fluence_map.size() != (size_t)arr[i].fmap_size.x * arr[i].fmap_size.y)
            throw std::invalid_argument("BEAM fluence_map does not hold fmap_size.x * fmap_size.y weights");
        fl[(size_t)i] = arr[i].fluence_map.data();
    }
}

}  // namespace

int initCudaConstandTex(SHM_DATA* datavols, MONO_KERNELS* mono, CONSTANTS* constants, int nrays, int deviceid,
                        int gpuid, bool, bool, bool) {
    const int k = device_slot(deviceid);
    if ((unsigned)nrays != constants->nphi * (constants->ntheta / 2))
        throw std::invalid_argument("nrays must equal nphi * ntheta / 2");
    cccs_problem p = problem_of(datavols, mono, constants);
    if (g_ctx[k]) {
        check(cccs_ctx_destroy(g_ctx[k]), "initCudaConstandTex");
        g_ctx[k] = nullptr;
    }
    check(cccs_ctx_create(&p, gpuid, &g_ctx[k]), "initCudaConstandTex");
    return 1;
}

#ifndef CCCS_SHIM_BEAMLET

int radconvolveTexture(MONO_KERNELS*, CONSTANTS* constants, BEAM* device_beam_arr, int nbeams, int, int deviceid,
                       int, bool, bool, bool) {
    cccs_ctx* c = ctx_of(deviceid);
    const size_t nv = (size_t)constants->size.x * constants->size.y * constants->size.z;
    float*& dose = g_dose[deviceid];
    if (!dose) {
        void* p = nullptr;
        check(cccs_dev_alloc(c, nv * sizeof(float), &p), "radconvolveTexture");
        dose = static_cast<float*>(p);
        check(cccs_dev_zero(c, dose, nv * sizeof(float)), "radconvolveTexture");
    }
    if (nbeams <= 0) return 1;
    std::vector<cccs_beam> beams;
    std::vector<const float*> fl;
    beams_of(device_beam_arr, nbeams, beams, fl);
    // the maps back to back on the device, in beam order
    size_t nf = 0;
    for (const cccs_beam& b : beams) nf += (size_t)b.fmap_size[0] * b.fmap_size[1];
    std::vector<float> host(nf);
    size_t o = 0;
    for (int i = 0; i < nbeams; i++) {
        const size_t n = (size_t)beams[(size_t)i].fmap_size[0] * beams[(size_t)i].fmap_size[1];
        std::memcpy(host.data() + o, fl[(size_t)i], n * sizeof(float));
        o += n;
    }
    void* d_fl = nullptr;
    check(cccs_dev_alloc(c, std::max<size_t>(1, nf) * sizeof(float), &d_fl), "radconvolveTexture");
    int rc = cccs_dev_upload(c, d_fl, host.data(), nf * sizeof(float));
    if (rc == CCCS_OK) rc = cccs_beam_dose_device(c, beams.data(), nbeams, static_cast<const float*>(d_fl), dose);
    if (rc == CCCS_OK) rc = cccs_ctx_sync(c);
    const std::string msg = rc == CCCS_OK ? std::string() : std::string(cccs_last_error());
    (void)cccs_dev_free(c, d_fl);
    if (rc != CCCS_OK) {
        if (rc == CCCS_EGEOM) throw ArraySizeError("radconvolveTexture: " + msg);
        throw std::runtime_error("radconvolveTexture: " + msg);
    }
    return 1;
}

int cuda_sum_device_float_dose(float* dose, int ndevices, CONSTANTS* constants, bool, bool) {
    const size_t nv = (size_t)constants->size.x * constants->size.y * constants->size.z;
    std::memset(dose, 0, nv * sizeof(float));
    std::vector<float> part(nv);
    for (int d = 0; d < ndevices; d++) {
        const int k = device_slot(d);
        if (!g_dose[k]) continue;
        check(cccs_dev_download(g_ctx[k], part.data(), g_dose[k], nv * sizeof(float)), "cuda_sum_device_float_dose");
        for (size_t v = 0; v < nv; v++) dose[v] += part[v];
    }
    return 1;
}

int freeCudaTexture(int ndevices, int*) {
    for (int d = 0; d < ndevices; d++) {
        const int k = device_slot(d);
        if (g_dose[k] && g_ctx[k]) (void)cccs_dev_free(g_ctx[k], g_dose[k]);
        g_dose[k] = nullptr;
        if (g_ctx[k]) check(cccs_ctx_destroy(g_ctx[k]), "freeCudaTexture");
        g_ctx[k] = nullptr;
    }
    return 1;
}

#else  // CCCS_SHIM_BEAMLET

namespace {
void check_h5(int rc, const char* where) {
    if (rc != CCCS_OK) throw std::runtime_error(std::string(where) + ": " + cccs_h5_last_error());
}

// BEAM::get_isocenter_type / get_isocenter_location / get_orientation_type (beam.cpp:36-89)
const char* iso_type_name(BEAM::ISO_T t) {
    switch (t) {
        case BEAM::ISO_T::UNSPEC: return "unspec";
        case BEAM::ISO_T::PTV_CENTROID: return "ptv";
        case BEAM::ISO_T::MANUAL: return "man";
    }
    return "unknown";
}
const char* iso_loc_name(BEAM::ISO_LOC_T t) {
    switch (t) {
        case BEAM::ISO_LOC_T::UNSPEC: return "unspec";
        case BEAM::ISO_LOC_T::IN_PTV: return "in";
        case BEAM::ISO_LOC_T::OUT_PTV: return "out";
    }
    return "unknown";
}
const char* orient_name(BEAM::ORIENT_T t) {
    switch (t) {
        case BEAM::ORIENT_T::DIRECTION: return "man";
        case BEAM::ORIENT_T::ANGLE: return "auto";
        default: return "unknown";
    }
}
}  // namespace

SparsifyManager::~SparsifyManager() {
    if (m_file) (void)cccs_h5_coeff_close(m_file);
}

void SparsifyManager::init(const std::string& h5filename, HEADER_PATIENT& h, unsigned int num_srworkers,
                           unsigned int, float sparsity_thresh, WRITE_STRATEGY write_strategy) {
    if (m_active) throw std::logic_error("Cannot initialize after activation of SparsifyManager");
    if (m_file) {
        (void)cccs_h5_coeff_close(m_file);
        m_file = nullptr;
    }
    m_thresh = sparsity_thresh;
    m_strategy = write_strategy;
    m_srworkers = num_srworkers;
    m_header = h;
    // the per-calculation header (_write_patient_metadata, sparse_io.cpp:210-326)
    cccs_problem p;
    std::memset(&p, 0, sizeof(p));
    p.size[0] = h.full_dicom_size.x; p.size[1] = h.full_dicom_size.y; p.size[2] = h.full_dicom_size.z;
    p.voxel[0] = h.voxel_size_cm.x; p.voxel[1] = h.voxel_size_cm.y; p.voxel[2] = h.voxel_size_cm.z;
    p.start[0] = h.dicom_start_cm.x; p.start[1] = h.dicom_start_cm.y; p.start[2] = h.dicom_start_cm.z;
    p.bbox_start[0] = h.bbox_start.x; p.bbox_start[1] = h.bbox_start.y; p.bbox_start[2] = h.bbox_start.z;
    p.bbox_size[0] = h.bbox_size.x; p.bbox_size[1] = h.bbox_size.y; p.bbox_size[2] = h.bbox_size.z;
    p.rev_latspacing = h.rev_latspacing_cm;
    p.rev_longspacing = h.rev_longspacing_cm;
    p.penumbra = h.penumbra_cm;
    p.kernel_extent = h.kernel_extent_cm;
    p.nphi = h.nphi;
    p.ntheta = h.ntheta;
    p.nradii = h.nradii;
    std::vector<const char*> names;
    for (const std::string& s : m_header.roi_order) names.push_back(s.c_str());
    cccs_h5_calc_specs specs;
    std::memset(&specs, 0, sizeof(specs));
    specs.version = nullptr;
    specs.nbeams = h.N_beams;
    specs.sparsity_thresh = sparsity_thresh;
    specs.beam_spectrum = m_header.beam_spectrum.c_str();
    specs.target_structure = m_header.target_structure.c_str();
    specs.nroi = (int)names.size();
    specs.roi_names = names.empty() ? nullptr : names.data();
    specs.roi_capacities = m_header.row_block_capacities.empty() ? nullptr : m_header.row_block_capacities.data();
    check_h5(cccs_h5_coeff_create(h5filename.c_str(), &p, &specs, -1, 50, &m_file), "SparsifyManager::init");
    // PER_BEAMLET: each column in its own sidecar file, linked from the central file
    // (sparsify_manager.cpp:123-174); PER_BEAM writes the central layout, as the reference does
    if (write_strategy == WRITE_STRATEGY::PER_BEAMLET)
        check_h5(cccs_h5_coeff_set_write_strategy(m_file, CCCS_H5_WRITE_PER_BEAMLET), "SparsifyManager::init");
}

void SparsifyManager::activate() {
    if (!m_file) throw std::logic_error("SparsifyManager::activate before init");
    m_active = true;
}

void SparsifyManager::deactivate() {
    m_active = false;
    if (m_file) {
        cccs_h5_coeff_file* f = m_file;
        m_file = nullptr;
        check_h5(cccs_h5_coeff_close(f), "SparsifyManager::deactivate");
    }
}

int radconvolveTexture(MONO_KERNELS*, CONSTANTS*, BEAM* device_beam_arr, int nbeams, int, int deviceid, int, int,
                       bool, bool, bool) {
    cccs_ctx* c = ctx_of(deviceid);
    cccs_h5_coeff_file* f = sparsifymanager.file();
    if (!f) throw std::logic_error("radconvolveTexture: sparsifymanager is not initialised");
    std::vector<cccs_beam> beams;
    std::vector<const float*> fl;
    beams_of(device_beam_arr, nbeams, beams, fl);
    // beams with consecutive uids share one request (the library batches their beamlets); the
    // columns go to the file under each BEAM's uid, as the reference's writer files them
    for (int i0 = 0; i0 < nbeams;) {
        int i1 = i0 + 1;
        while (i1 < nbeams && device_beam_arr[i1].uid == device_beam_arr[i0].uid + (uint)(i1 - i0)) i1++;
        std::vector<cccs_h5_beam_labels> labels((size_t)(i1 - i0));
        for (int i = i0; i < i1; i++) {
            labels[(size_t)(i - i0)].isocenter_type = iso_type_name(device_beam_arr[i].isocenter_type);
            labels[(size_t)(i - i0)].isocenter_location = iso_loc_name(device_beam_arr[i].isocenter_location);
            labels[(size_t)(i - i0)].orientation_type = orient_name(device_beam_arr[i].orientation_type);
        }
        check_h5(cccs_h5_coeff_add_beams(f, beams.data() + i0, i1 - i0, fl.data() + i0, (int)device_beam_arr[i0].uid,
                                         labels.data()),
                 "radconvolveTexture");
        const int rc = cccs_beamlet_dose(c, beams.data() + i0, i1 - i0, fl.data() + i0, sparsifymanager.threshold(),
                                         CCCS_BEAMLET_FAITHFUL, cccs_h5_coeff_sink, f);
        if (rc != CCCS_OK) {
            const std::string detail = cccs_h5_last_error();
            check(rc, detail.empty() ? "radconvolveTexture" : ("radconvolveTexture (" + detail + ")").c_str());
        }
        i0 = i1;
    }
    return 1;
}

int freeCudaTexture(int ndevices, int*, bool) {
    for (int d = 0; d < ndevices; d++) {
        const int k = device_slot(d);
        if (g_ctx[k]) check(cccs_ctx_destroy(g_ctx[k]), "freeCudaTexture");
        g_ctx[k] = nullptr;
    }
    return 1;
}

#endif
