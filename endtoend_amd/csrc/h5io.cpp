// h5io.cpp -- the HDF5 file layer (libcccs_h5.so): Dose.h5 volumes, Dose_Coefficients.h5
// coefficient files, fluence-map and ROI-list inputs, over the HDF5 C library.  Host code only.
//
// The group names, attribute names, file types and layouts are the reference's on-disk format
// (DoseCalcIO/volume.h:118-182, sparse_io.cpp:178-326, beam.cpp:134-298, fmaps_io.cpp:10-184,
// roi.cpp:16-185, io_data_structs.h:80-134, dosecalc-beamlet/sparsify_manager.cpp:11-221), kept
// identical so that the reference's readers (read_fmaps, ROIMaskList::readFromFile,
// extern/python/read_dose_data.py) read these files unchanged.
#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "cccs_h5.h"

#include <errno.h>
#include <ftw.h>
#include <sys/stat.h>
#include <unistd.h>

namespace {

thread_local std::string g_err;
std::mutex g_lock;  // the HDF5 library build is not assumed thread-safe

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

// owning handle of one HDF5 identifier
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id() = default;
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    Id(Id&& o) noexcept : id(o.id), close(o.close) { o.id = -1; }
    ~Id() {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
    operator hid_t() const { return id; }
};

Id group_create(hid_t loc, const char* name) { return Id(H5Gcreate2(loc, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose); }
Id group_open(hid_t loc, const char* name) { return Id(H5Gopen2(loc, name, H5P_DEFAULT), H5Gclose); }
Id scalar_space() { return Id(H5Screate(H5S_SCALAR), H5Sclose); }
Id simple_space(int rank, const hsize_t* dims) { return Id(H5Screate_simple(rank, dims, nullptr), H5Sclose); }
Id array_type(hid_t base, hsize_t n) { return Id(H5Tarray_create2(base, 1, &n), H5Tclose); }

// one attribute of the given file type over `space`, written from memory type `mem`
bool attr_write(hid_t loc, const char* name, hid_t ftype, hid_t space, hid_t mem, const void* buf) {
    Id a(H5Acreate2(loc, name, ftype, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Awrite(a, mem, buf) >= 0;
}
bool attr_scalar(hid_t loc, const char* name, hid_t ftype, hid_t mem, const void* buf) {
    Id s = scalar_space();
    return attr_write(loc, name, ftype, s, mem, buf);
}
// a 3-tuple stored as one ARRAY-typed scalar (the calc_specs and ArrayProps tuples)
bool attr_tuple3(hid_t loc, const char* name, hid_t fbase, hid_t mbase, const void* buf) {
    Id ft = array_type(fbase, 3), mt = array_type(mbase, 3), s = scalar_space();
    return ft.ok() && mt.ok() && attr_write(loc, name, ft, s, mt, buf);
}
// an n-vector stored over a 1-D dataspace (the Volume tuples, row-block lists)
bool attr_vector(hid_t loc, const char* name, hid_t ftype, hid_t mem, hsize_t n, const void* buf) {
    Id s = simple_space(1, &n);
    return s.ok() && attr_write(loc, name, ftype, s, mem, buf);
}
// fixed-length, NUL-terminated C string of strlen + 1 bytes
bool attr_string(hid_t loc, const char* name, const std::string& v) {
    Id t(H5Tcopy(H5T_C_S1), H5Tclose);
    if (!t.ok() || H5Tset_size(t, v.size() + 1) < 0) return false;
    return attr_scalar(loc, name, t, t, v.c_str());
}
bool attr_vlen_strings(hid_t loc, const char* name, const std::vector<const char*>& v) {
    Id t(H5Tcopy(H5T_C_S1), H5Tclose);
    if (!t.ok() || H5Tset_size(t, H5T_VARIABLE) < 0) return false;
    return attr_vector(loc, name, t, t, v.size(), v.data());
}

bool write_file_version(hid_t file, unsigned magic, unsigned major, unsigned minor) {
    Id g = group_create(file, "filetype");
    return g.ok() && attr_scalar(g, "ftmagic", H5T_STD_U8LE, H5T_NATIVE_UINT, &magic) &&
           attr_scalar(g, "ftversionmajor", H5T_STD_U8LE, H5T_NATIVE_UINT, &major) &&
           attr_scalar(g, "ftversionminor", H5T_STD_U8LE, H5T_NATIVE_UINT, &minor);
}

// string attribute of any fixed or variable string type, read as std::string
bool attr_read_string(hid_t loc, const char* name, std::string& out) {
    Id a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return false;
    Id t(H5Aget_type(a), H5Tclose);
    if (!t.ok() || H5Tget_class(t) != H5T_STRING) return false;
    if (H5Tis_variable_str(t) > 0) {
        char* p = nullptr;
        Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(mt, H5T_VARIABLE);
        if (H5Aread(a, mt, &p) < 0) return false;
        out = p ? p : "";
        Id s = scalar_space();
        H5Dvlen_reclaim(mt, s, H5P_DEFAULT, &p);
        return true;
    }
    const size_t n = H5Tget_size(t);
    std::vector<char> buf(n + 1, 0);
    if (H5Aread(a, t, buf.data()) < 0) return false;
    out.assign(buf.data(), strnlen(buf.data(), n));
    return true;
}
bool attr_read(hid_t loc, const char* name, hid_t mem, void* buf) {
    Id a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Aread(a, mem, buf) >= 0;
}
bool attr_read_tuple3(hid_t loc, const char* name, hid_t mbase, void* buf) {
    Id a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return false;
    Id t(H5Aget_type(a), H5Tclose);
    if (!t.ok()) return false;
    if (H5Tget_class(t) == H5T_ARRAY) {  // ARRAY-typed scalar
        Id mt = array_type(mbase, 3);
        return H5Aread(a, mt, buf) >= 0;
    }
    Id s(H5Aget_space(a), H5Sclose);  // or a 3-vector
    if (H5Sget_simple_extent_npoints(s) != 3) return false;
    return H5Aread(a, mbase, buf) >= 0;
}
hssize_t attr_npoints(hid_t loc, const char* name) {
    Id a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return -1;
    Id s(H5Aget_space(a), H5Sclose);
    return H5Sget_simple_extent_npoints(s);
}

// group names below `loc`, in name order
herr_t collect_name(hid_t, const char* name, const H5L_info_t*, void* op) {
    static_cast<std::vector<std::string>*>(op)->push_back(name);
    return 0;
}
bool list_links(hid_t loc, std::vector<std::string>& names) {
    hsize_t idx = 0;
    return H5Literate(loc, H5_INDEX_NAME, H5_ITER_INC, &idx, collect_name, &names) >= 0;
}

// BEAM's compound record (beam.h:81-93): member order, types and the natural-alignment offsets
// of that struct, which are also the file offsets (beam.cpp:134-163)
struct BeamRecord {
    uint16_t uid;
    float gantry_rot_rad;
    float couch_rot_rad;
    float coll_rot_rad;
    float src_coords_cm[3];
    float direction[3];
    float iso_coords_cm[3];
    uint32_t fmap_dims[2];
    float beamlet_size[2];
};
static_assert(sizeof(BeamRecord) == 68, "BEAM record layout");

Id beam_record_type(bool file) {
    Id t(H5Tcreate(H5T_COMPOUND, sizeof(BeamRecord)), H5Tclose);
    const hid_t f32 = file ? H5T_IEEE_F32LE : H5T_NATIVE_FLOAT;
    const hid_t u32 = file ? H5T_STD_U32LE : H5T_NATIVE_UINT;
    const hid_t u16 = file ? H5T_STD_U16LE : H5T_NATIVE_USHORT;
    Id v3 = array_type(f32, 3), u2 = array_type(u32, 2), f2 = array_type(f32, 2);
    H5Tinsert(t, "uid", offsetof(BeamRecord, uid), u16);
    H5Tinsert(t, "gantry_rot_rad", offsetof(BeamRecord, gantry_rot_rad), f32);
    H5Tinsert(t, "couch_rot_rad", offsetof(BeamRecord, couch_rot_rad), f32);
    H5Tinsert(t, "coll_rot_rad", offsetof(BeamRecord, coll_rot_rad), f32);
    H5Tinsert(t, "src_coords_cm", offsetof(BeamRecord, src_coords_cm), v3);
    H5Tinsert(t, "direction", offsetof(BeamRecord, direction), v3);
    H5Tinsert(t, "iso_coords_cm", offsetof(BeamRecord, iso_coords_cm), v3);
    H5Tinsert(t, "fmap_dims", offsetof(BeamRecord, fmap_dims), u2);
    H5Tinsert(t, "beamlet_size_cm", offsetof(BeamRecord, beamlet_size), f2);
    return t;
}

std::string beam_group_name(const char* parent, int uid) {
    char b[64];
    std::snprintf(b, sizeof(b), "%s/beam_%05d", parent, uid);
    return b;
}

// the attributes of one beam's metadata group (_write_beam_metadata + BEAM::_writeToHDF5,
// sparse_io.cpp:248-262, beam.cpp:248-298)
bool write_beam_attrs(hid_t g, int uid, unsigned nbeamlets, const cccs_beam& b, const float* fluence,
                      const cccs_h5_beam_labels* L) {
    Id ftype = beam_record_type(true), mtype = beam_record_type(false);
    const unsigned u = (unsigned)uid;
    BeamRecord r{};
    r.uid = (uint16_t)uid;
    r.gantry_rot_rad = b.azimuth;
    r.couch_rot_rad = b.zenith;
    r.coll_rot_rad = b.coll;
    for (int a = 0; a < 3; a++) {
        r.src_coords_cm[a] = b.source[a];
        r.direction[a] = b.direction[a];
        r.iso_coords_cm[a] = b.isocenter[a];
    }
    r.fmap_dims[0] = b.fmap_size[0];
    r.fmap_dims[1] = b.fmap_size[1];
    r.beamlet_size[0] = b.beamlet_size[0];
    r.beamlet_size[1] = b.beamlet_size[1];
    const hsize_t fdims[2] = {b.fmap_size[0], b.fmap_size[1]};
    Id fs = simple_space(2, fdims);
    return attr_scalar(g, "beam_uid", H5T_STD_U16LE, H5T_NATIVE_UINT, &u) &&
           attr_scalar(g, "N_beamlets", H5T_STD_U16LE, H5T_NATIVE_UINT, &nbeamlets) &&
           attr_string(g, "isocenter_type", L && L->isocenter_type ? L->isocenter_type : "unspec") &&
           attr_string(g, "isocenter_location", L && L->isocenter_location ? L->isocenter_location : "unspec") &&
           attr_string(g, "orientation_type", L && L->orientation_type ? L->orientation_type : "unknown") &&
           attr_write(g, "fmap_weights", H5T_IEEE_F32LE, fs, H5T_NATIVE_FLOAT, fluence) &&
           attr_scalar(g, "beam_specs", ftype, mtype, &r);
}

// the attributes of one beamlet's group (_write_beamlet_metadata, sparse_io.cpp:209-245)
bool write_beamlet_attrs(hid_t g, const cccs_sparse_column* c) {
    const unsigned bu = (unsigned)c->beamlet_index;
    const uint64_t n = c->nnz;
    bool ok = attr_scalar(g, "beamlet_uid", H5T_STD_U16LE, H5T_NATIVE_UINT, &bu) &&
              attr_scalar(g, "N_coeffs", H5T_STD_U64LE, H5T_NATIVE_UINT64, &n) &&
              attr_scalar(g, "perc_nonzero", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &c->perc_nonzero) &&
              attr_scalar(g, "perc_dropped", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &c->perc_dropped);
    if (ok && c->nblocks > 0)
        ok = attr_vector(g, "row_block_sizes", H5T_STD_U64LE, H5T_NATIVE_UINT64, c->nblocks, c->row_block_sizes) &&
             attr_scalar(g, "perc_reducedrop", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &c->perc_reducedrop);
    return ok;
}

// the lindex / coeffs datasets of one beamlet (_write_beamlet_to_hdf5, sparse_io.cpp:178-207)
bool write_beamlet_data(hid_t g, const cccs_sparse_column* c, int compress_lvl, uint64_t chunksize) {
    const uint64_t n = c->nnz;
    Id plist(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    if (compress_lvl >= 0) {
        const hsize_t chunk = std::min<hsize_t>(n, chunksize);
        if (H5Pset_chunk(plist, 1, &chunk) < 0 || H5Pset_deflate(plist, (unsigned)std::min(9, compress_lvl)) < 0)
            return false;
    }
    const hsize_t dims = n;
    Id s = simple_space(1, &dims);
    Id k(H5Dcreate2(g, "lindex", H5T_STD_U64LE, s, H5P_DEFAULT, plist, H5P_DEFAULT), H5Dclose);
    if (!k.ok() || H5Dwrite(k, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, c->keys) < 0) return false;
    Id v(H5Dcreate2(g, "coeffs", H5T_IEEE_F32LE, s, H5P_DEFAULT, plist, H5P_DEFAULT), H5Dclose);
    return v.ok() && H5Dwrite(v, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, c->values) >= 0;
}

int rm_entry(const char* p, const struct stat*, int, struct FTW*) { return ::remove(p); }

// removes the directory tree at `dir` (no symlinks followed); true when nothing is left
bool remove_tree(const std::string& dir) {
    struct stat st;
    if (::lstat(dir.c_str(), &st) != 0) return true;
    return ::nftw(dir.c_str(), rm_entry, 16, FTW_DEPTH | FTW_PHYS) == 0;
}

bool make_dir(const std::string& dir) { return ::mkdir(dir.c_str(), 0755) == 0 || errno == EEXIST; }

// last path component, trailing separators ignored (POSIX basename)
std::string base_name(std::string p) {
    while (p.size() > 1 && p.back() == '/') p.pop_back();
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}

}  // namespace

struct cccs_h5_coeff_file {
    hid_t file = -1;
    std::string path;
    int strategy = CCCS_H5_WRITE_CENTRAL;
    std::string sidecar;  // PER_BEAMLET: "<path>_external/"
    int compress_lvl = -1;
    uint64_t chunksize = 50;
    std::vector<int> request_uids;  // request beam index -> beam uid
    std::set<int> beams_written;
    uint64_t columns = 0;
    uint64_t coeffs = 0;
};

extern "C" {

const char* cccs_h5_last_error(void) { return g_err.c_str(); }

int cccs_h5_write_volume(const char* path, const char* dset_name, const float* data, const uint32_t size[3],
                         const float start[3], const float voxel[3]) {
    g_err.clear();
    if (!path || !data || !size || !start || !voxel) return fail(CCCS_EINVAL, "null argument");
    if (!size[0] || !size[1] || !size[2]) return fail(CCCS_EINVAL, "empty volume");
    std::lock_guard<std::mutex> lk(g_lock);
    quiet();
    Id f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(CCCS_EIO, std::string("cannot create ") + path);
    if (!write_file_version(f, 0x2A, 1, 0)) return fail(CCCS_EIO, "cannot write /filetype");
    const hsize_t dims[3] = {size[2], size[1], size[0]};
    Id s = simple_space(3, dims);
    Id d(H5Dcreate2(f, dset_name ? dset_name : "volume", H5T_IEEE_F32LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
         H5Dclose);
    if (!d.ok() || H5Dwrite(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        return fail(CCCS_EIO, "cannot write the volume dataset");
    if (!attr_vector(d, "dicom_start_cm", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, 3, start) ||
        !attr_vector(d, "voxel_size_cm", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, 3, voxel))
        return fail(CCCS_EIO, "cannot write the volume attributes");
    if (H5Fflush(f, H5F_SCOPE_LOCAL) < 0) return fail(CCCS_EIO, "flush failed");
    return CCCS_OK;
}

int cccs_h5_read_volume(const char* path, const char* dset_name, float* out, size_t cap, uint32_t size[3],
                        float start[3], float voxel[3]) {
    g_err.clear();
    if (!path || !size || !start || !voxel) return fail(CCCS_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(g_lock);
    quiet();
    Id f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(CCCS_EIO, std::string("cannot open ") + path);
    const char* name = dset_name ? dset_name : "data";
    Id d(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(CCCS_EIO, std::string("no dataset ") + name);
    Id s(H5Dget_space(d), H5Sclose);
    hsize_t dims[3] = {0, 0, 0};
    if (H5Sget_simple_extent_ndims(s) != 3 || H5Sget_simple_extent_dims(s, dims, nullptr) != 3)
        return fail(CCCS_EIO, "volume dataset is not 3-D");
    if (dims[0] > UINT32_MAX || dims[1] > UINT32_MAX || dims[2] > UINT32_MAX) return fail(CCCS_EIO, "volume too large");
    size[0] = (uint32_t)dims[2];
    size[1] = (uint32_t)dims[1];
    size[2] = (uint32_t)dims[0];
    if (!attr_read(d, "dicom_start_cm", H5T_NATIVE_FLOAT, start) || !attr_read(d, "voxel_size_cm", H5T_NATIVE_FLOAT, voxel))
        return fail(CCCS_EIO, "missing dicom_start_cm / voxel_size_cm");
    if (!out) return CCCS_OK;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    if (cap < n) return fail(CCCS_EINVAL, "output buffer too small");
    if (H5Dread(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        return fail(CCCS_EIO, "cannot read the volume dataset");
    return CCCS_OK;
}

int cccs_h5_coeff_create(const char* path, const cccs_problem* prob, const cccs_h5_calc_specs* specs,
                         int compress_lvl, uint64_t chunksize, cccs_h5_coeff_file** out) {
    g_err.clear();
    if (!path || !prob || !specs || !out) return fail(CCCS_EINVAL, "null argument");
    *out = nullptr;
    if (specs->nbeams < 0 || specs->nbeams > 65535) return fail(CCCS_EINVAL, "N_beams must fit 16 bits");
    for (int a = 0; a < 3; a++)
        if (prob->size[a] > 65535 || prob->bbox_start[a] > 65535 || prob->bbox_size[a] > 65535)
            return fail(CCCS_EINVAL, "volume and calc_bbox sizes must fit 16 bits");
    if (specs->nroi < 0 || (specs->nroi > 0 && (!specs->roi_names || !specs->roi_capacities)))
        return fail(CCCS_EINVAL, "ROI names and capacities required");
    if (compress_lvl >= 0 && chunksize == 0) return fail(CCCS_EINVAL, "chunksize must be positive");
    std::lock_guard<std::mutex> lk(g_lock);
    quiet();
    hid_t fid = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (fid < 0) return fail(CCCS_EIO, std::string("cannot create ") + path);
    auto abort_create = [&](const char* msg) {
        H5Fclose(fid);
        return fail(CCCS_EIO, msg);
    };
    if (!write_file_version(fid, 0x2B, 1, 2)) return abort_create("cannot write /filetype");
    {
        Id g = group_create(fid, "calc_specs");
        if (!g.ok()) return abort_create("cannot create /calc_specs");
        const unsigned nbeams = (unsigned)specs->nbeams, nphi = prob->nphi, ntheta = prob->ntheta, nradii = prob->nradii;
        const unsigned dsize[3] = {prob->size[0], prob->size[1], prob->size[2]};
        const unsigned bstart[3] = {prob->bbox_start[0], prob->bbox_start[1], prob->bbox_start[2]};
        const unsigned bsize[3] = {prob->bbox_size[0], prob->bbox_size[1], prob->bbox_size[2]};
        bool ok = attr_string(g, "dosecalc_version", specs->version ? specs->version : "") &&
                  attr_scalar(g, "N_beams", H5T_STD_U16LE, H5T_NATIVE_UINT, &nbeams) &&
                  attr_tuple3(g, "dicom_start_cm", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, prob->start) &&
                  attr_tuple3(g, "full_dicom_size", H5T_STD_U16LE, H5T_NATIVE_UINT, dsize) &&
                  attr_tuple3(g, "voxel_size_cm", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, prob->voxel) &&
                  attr_scalar(g, "convlat_cm", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &prob->rev_latspacing) &&
                  attr_scalar(g, "convstep_cm", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &prob->rev_longspacing) &&
                  attr_tuple3(g, "calc_bbox_start", H5T_STD_U16LE, H5T_NATIVE_UINT, bstart) &&
                  attr_tuple3(g, "calc_bbox_size", H5T_STD_U16LE, H5T_NATIVE_UINT, bsize) &&
                  attr_scalar(g, "penumbra_cm", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &prob->penumbra) &&
                  attr_scalar(g, "sparsity_thresh", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &specs->sparsity_thresh) &&
                  attr_scalar(g, "nphi", H5T_STD_U16LE, H5T_NATIVE_UINT, &nphi) &&
                  attr_scalar(g, "ntheta", H5T_STD_U16LE, H5T_NATIVE_UINT, &ntheta) &&
                  attr_scalar(g, "nradii", H5T_STD_U16LE, H5T_NATIVE_UINT, &nradii) &&
                  attr_scalar(g, "kernel_extent_cm", H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &prob->kernel_extent) &&
                  attr_string(g, "beam_spectrum", specs->beam_spectrum ? specs->beam_spectrum : "") &&
                  attr_string(g, "target_structure", specs->target_structure ? specs->target_structure : "");
        if (ok && specs->nroi > 0) {
            std::vector<const char*> names(specs->roi_names, specs->roi_names + specs->nroi);
            for (const char* n : names)
                if (!n) return abort_create("null ROI name");
            ok = attr_vector(g, "row_block_capacities", H5T_STD_U64LE, H5T_NATIVE_UINT64, specs->nroi,
                             specs->roi_capacities) &&
                 attr_vlen_strings(g, "roi_order", names);
        }
        if (!ok) return abort_create("cannot write /calc_specs attributes");
    }
    {
        Id a = group_create(fid, "/beams"), b = group_create(fid, "/beams/metadata"), c = group_create(fid, "/beams/data");
        if (!a.ok() || !b.ok() || !c.ok()) return abort_create("cannot create the /beams groups");
    }
    auto* h = new cccs_h5_coeff_file();
    h->file = fid;
    h->path = path;
    h->compress_lvl = compress_lvl;
    h->chunksize = chunksize;
    *out = h;
    return CCCS_OK;
}

int cccs_h5_coeff_set_write_strategy(cccs_h5_coeff_file* f, int strategy) {
    g_err.clear();
    if (!f) return fail(CCCS_EINVAL, "null argument");
    if (strategy != CCCS_H5_WRITE_CENTRAL && strategy != CCCS_H5_WRITE_PER_BEAMLET)
        return fail(CCCS_EINVAL, "unknown write strategy");
    std::lock_guard<std::mutex> lk(g_lock);
    if (!f->beams_written.empty() || f->columns) return fail(CCCS_EINVAL, "set the write strategy before any beam");
    f->strategy = strategy;
    f->sidecar.clear();
    if (strategy == CCCS_H5_WRITE_PER_BEAMLET) {
        // SparsifyManager::init (sparsify_manager.cpp:196-204): a fresh "<file>_external/" directory
        f->sidecar = f->path + "_external/";
        if (!remove_tree(f->sidecar) || !make_dir(f->sidecar)) return fail(CCCS_EIO, "cannot create " + f->sidecar);
    }
    return CCCS_OK;
}

int cccs_h5_coeff_add_beams(cccs_h5_coeff_file* f, const cccs_beam* beams, int nbeams, const float* const* fluence,
                            int uid_base, const cccs_h5_beam_labels* labels) {
    g_err.clear();
    if (!f || !beams || nbeams < 0 || !fluence) return fail(CCCS_EINVAL, "null argument");
    if (uid_base < 0 || uid_base + nbeams - 1 > 65535) return fail(CCCS_EINVAL, "beam uids must fit 16 bits");
    for (int i = 0; i < nbeams; i++) {
        if (!fluence[i]) return fail(CCCS_EINVAL, "null fluence map");
        if ((uint64_t)beams[i].fmap_size[0] * beams[i].fmap_size[1] == 0) return fail(CCCS_EINVAL, "empty fluence map");
    }
    std::lock_guard<std::mutex> lk(g_lock);
    quiet();
    f->request_uids.assign(nbeams, -1);
    for (int i = 0; i < nbeams; i++) {
        const int uid = uid_base + i;
        f->request_uids[i] = uid;
        const cccs_beam& b = beams[i];
        const size_t n = (size_t)b.fmap_size[0] * b.fmap_size[1];
        size_t active = 0;
        for (size_t k = 0; k < n; k++) active += fluence[i][k] > 0.0f;
        // the reference opens a beam's groups on its first beamlet, so a beam with no active
        // beamlet leaves no trace in the file
        if (active == 0 || f->beams_written.count(uid)) continue;
        if (active > 65535) return fail(CCCS_EINVAL, "N_beamlets must fit 16 bits");
        const std::string meta = beam_group_name("/beams/metadata", uid), data = beam_group_name("/beams/data", uid);
        Id dg = group_create(f->file, data.c_str());
        Id g = group_create(f->file, meta.c_str());
        if (!dg.ok() || !g.ok()) return fail(CCCS_EIO, "cannot create the groups of " + meta);
        if (!write_beam_attrs(g, uid, (unsigned)active, b, fluence[i], labels ? &labels[i] : nullptr))
            return fail(CCCS_EIO, "cannot write the attributes of " + meta);
        if (f->strategy == CCCS_H5_WRITE_PER_BEAMLET) {
            char sub[32];
            std::snprintf(sub, sizeof(sub), "beam_%05d", uid);
            if (!make_dir(f->sidecar + sub)) return fail(CCCS_EIO, "cannot create " + f->sidecar + sub);
        }
        f->beams_written.insert(uid);
    }
    return CCCS_OK;
}

int cccs_h5_coeff_add_column(cccs_h5_coeff_file* f, const cccs_sparse_column* c) {
    g_err.clear();
    if (!f || !c) return fail(CCCS_EINVAL, "null argument");
    if (c->beam_index < 0) return fail(CCCS_EINVAL, "negative beam index");
    {
        std::lock_guard<std::mutex> lk(g_lock);
        if (c->beam_index >= (int)f->request_uids.size())
            return fail(CCCS_EINVAL, "column of an undeclared beam (call cccs_h5_coeff_add_beams first)");
    }
    if (c->beamlet_index < 0 || c->beamlet_index > 65535) return fail(CCCS_EINVAL, "beamlet uid must fit 16 bits");
    if (c->nnz && (!c->keys || !c->values)) return fail(CCCS_EINVAL, "null column data");
    if (c->nblocks < 0 || (c->nblocks > 0 && !c->row_block_sizes)) return fail(CCCS_EINVAL, "null row-block sizes");
    if (c->nnz == 0) return CCCS_OK;  // the reference writes no group for an empty column
    std::lock_guard<std::mutex> lk(g_lock);
    quiet();
    if (c->beam_index >= (int)f->request_uids.size()) return fail(CCCS_EINVAL, "column of an undeclared beam");
    const int uid = f->request_uids[c->beam_index];
    if (!f->beams_written.count(uid)) return fail(CCCS_EINVAL, "column of a beam with no active beamlet");
    char name[96];
    std::snprintf(name, sizeof(name), "/beams/data/beam_%05d/beamlet_%05d", uid, c->beamlet_index);
    const uint64_t n = c->nnz;
    if (f->strategy == CCCS_H5_WRITE_PER_BEAMLET) {
        // the column goes to its own sidecar file (metadata + datasets at its root); the central
        // group keeps the metadata and external links to the sidecar's datasets
        // (W2FWorker_PerBeamlet::write_beamlet, sparsify_manager.cpp:129-174)
        char sub[64];
        std::snprintf(sub, sizeof(sub), "beam_%05d/beam%05d_beamlet%05d.bmlt", uid, uid, c->beamlet_index);
        const std::string file = f->sidecar + sub;
        {
            Id sf(H5Fcreate(file.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
            if (!sf.ok()) return fail(CCCS_EIO, "cannot create " + file);
            Id root(H5Gopen2(sf, "/", H5P_DEFAULT), H5Gclose);
            if (!root.ok() || !write_beamlet_attrs(root, c) || !write_beamlet_data(root, c, f->compress_lvl, f->chunksize))
                return fail(CCCS_EIO, "cannot write " + file);
        }
        Id g = group_create(f->file, name);
        if (!g.ok()) return fail(CCCS_EIO, std::string("cannot create ") + name + " (written twice?)");
        if (!write_beamlet_attrs(g, c)) return fail(CCCS_EIO, std::string("cannot write the attributes of ") + name);
        const std::string target = base_name(f->sidecar) + "/" + sub;
        if (H5Lcreate_external(target.c_str(), "/lindex", g, "lindex", H5P_DEFAULT, H5P_DEFAULT) < 0 ||
            H5Lcreate_external(target.c_str(), "/coeffs", g, "coeffs", H5P_DEFAULT, H5P_DEFAULT) < 0)
            return fail(CCCS_EIO, std::string("cannot link ") + name + " to " + target);
    } else {
        Id g = group_create(f->file, name);
        if (!g.ok()) return fail(CCCS_EIO, std::string("cannot create ") + name + " (written twice?)");
        if (!write_beamlet_attrs(g, c)) return fail(CCCS_EIO, std::string("cannot write the attributes of ") + name);
        if (!write_beamlet_data(g, c, f->compress_lvl, f->chunksize))
            return fail(CCCS_EIO, std::string("cannot write the datasets of ") + name);
    }
    f->columns++;
    f->coeffs += n;
    return CCCS_OK;
}

int cccs_h5_coeff_sink(void* user, const cccs_sparse_column* column) {
    return cccs_h5_coeff_add_column(static_cast<cccs_h5_coeff_file*>(user), column) == CCCS_OK ? 0 : 1;
}

int cccs_h5_coeff_counts(const cccs_h5_coeff_file* f, uint64_t* columns, uint64_t* coeffs) {
    if (!f || !columns || !coeffs) return fail(CCCS_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(g_lock);
    *columns = f->columns;
    *coeffs = f->coeffs;
    return CCCS_OK;
}

int cccs_h5_coeff_close(cccs_h5_coeff_file* f) {
    g_err.clear();
    if (!f) return fail(CCCS_EINVAL, "null argument");
    int st = CCCS_OK;
    {
        std::lock_guard<std::mutex> lk(g_lock);
        quiet();
        if (f->file >= 0 && (H5Fflush(f->file, H5F_SCOPE_LOCAL) < 0 || H5Fclose(f->file) < 0))
            st = fail(CCCS_EIO, "cannot flush / close the coefficient file");
    }
    delete f;
    return st;
}

int cccs_h5_write_fmaps(const char* path, const cccs_beam* beams, int nbeams, const float* const* fluence,
                        const cccs_h5_beam_labels* labels) {
    g_err.clear();
    if (!path || nbeams < 0 || (nbeams && (!beams || !fluence))) return fail(CCCS_EINVAL, "null argument");
    if (nbeams > 65536) return fail(CCCS_EINVAL, "beam uids must fit 16 bits");
    std::vector<unsigned> nz(nbeams, 0);
    for (int i = 0; i < nbeams; i++) {
        if (!fluence[i]) return fail(CCCS_EINVAL, "null fluence map");
        const size_t n = (size_t)beams[i].fmap_size[0] * beams[i].fmap_size[1];
        size_t c = 0;
        for (size_t k = 0; k < n; k++) c += fluence[i][k] != 0.0f;  // count_nonzero (fmaps.py:94)
        if (c > 65535) return fail(CCCS_EINVAL, "N_beamlets must fit 16 bits");
        nz[i] = (unsigned)c;
    }
    static const cccs_h5_beam_labels kFmapsLabels = {"man", "unspec", "auto"};  // fmaps.py:97-99
    std::lock_guard<std::mutex> lk(g_lock);
    quiet();
    Id f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(CCCS_EIO, std::string("cannot create ") + path);
    if (!write_file_version(f, 0x2C, 1, 1)) return fail(CCCS_EIO, "cannot write /filetype");
    Id a = group_create(f, "/beams"), m = group_create(f, "/beams/metadata");
    if (!a.ok() || !m.ok()) return fail(CCCS_EIO, "cannot create the /beams groups");
    for (int i = 0; i < nbeams; i++) {
        const std::string meta = beam_group_name("/beams/metadata", i);
        Id g = group_create(f, meta.c_str());
        if (!g.ok() || !write_beam_attrs(g, i, nz[i], beams[i], fluence[i], labels ? &labels[i] : &kFmapsLabels))
            return fail(CCCS_EIO, "cannot write " + meta);
    }
    if (H5Fflush(f, H5F_SCOPE_LOCAL) < 0) return fail(CCCS_EIO, "flush failed");
    return CCCS_OK;
}

int cccs_h5_read_fmaps(const char* path, cccs_beam* beams, int* uids, int cap_beams, int* nbeams, float* weights,
                       size_t cap_weights, size_t* nweights) {
    g_err.clear();
    if (!path || !nbeams || !nweights) return fail(CCCS_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(g_lock);
    quiet();
    Id f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(CCCS_EIO, std::string("cannot open ") + path);
    // the coefficient format keeps the beam records under /beams/metadata, the optimiser's
    // fluence files directly under /beams (fmaps_io.cpp:150-155)
    Id g = H5Lexists(f, "/beams/metadata", H5P_DEFAULT) > 0 ? group_open(f, "/beams/metadata") : group_open(f, "/beams");
    if (!g.ok()) return fail(CCCS_EIO, "no /beams group");
    std::vector<std::string> names;
    if (!list_links(g, names)) return fail(CCCS_EIO, "cannot list the beam groups");
    struct Entry {
        unsigned uid;
        cccs_beam beam;
        std::vector<float> w;
    };
    std::vector<Entry> all;
    Id mtype = beam_record_type(false);
    for (const std::string& nm : names) {
        Id bg = group_open(g, nm.c_str());
        if (!bg.ok()) return fail(CCCS_EIO, "cannot open beam group " + nm);
        Entry e{};
        BeamRecord r{};
        if (!attr_read(bg, "beam_specs", mtype, &r)) return fail(CCCS_EIO, "beam_specs missing in " + nm);
        if (!attr_read(bg, "beam_uid", H5T_NATIVE_UINT, &e.uid)) return fail(CCCS_EIO, "beam_uid missing in " + nm);
        cccs_beam& b = e.beam;
        b.azimuth = r.gantry_rot_rad;
        b.zenith = r.couch_rot_rad;
        b.coll = r.coll_rot_rad;
        float d2 = 0.0f;
        for (int a = 0; a < 3; a++) {
            b.source[a] = r.src_coords_cm[a];
            b.direction[a] = r.direction[a];
            b.isocenter[a] = r.iso_coords_cm[a];
            const float t = r.iso_coords_cm[a] - r.src_coords_cm[a];
            d2 += t * t;
        }
        b.sad = std::sqrt(d2);  // beam.cpp:217: sad = |isocenter - source|
        b.fmap_size[0] = r.fmap_dims[0];
        b.fmap_size[1] = r.fmap_dims[1];
        b.beamlet_size[0] = r.beamlet_size[0];
        b.beamlet_size[1] = r.beamlet_size[1];
        const hssize_t nw = attr_npoints(bg, "fmap_weights");
        if (nw < 0) return fail(CCCS_EIO, "fmap_weights missing in " + nm);
        e.w.resize((size_t)nw);
        if (nw && !attr_read(bg, "fmap_weights", H5T_NATIVE_FLOAT, e.w.data()))
            return fail(CCCS_EIO, "cannot read fmap_weights of " + nm);
        all.push_back(std::move(e));
    }
    std::stable_sort(all.begin(), all.end(), [](const Entry& a, const Entry& b) { return a.uid < b.uid; });
    size_t total = 0;
    for (const Entry& e : all) total += e.w.size();
    *nbeams = (int)all.size();
    *nweights = total;
    if (!beams) return CCCS_OK;
    if (cap_beams < (int)all.size() || (total && (!weights || cap_weights < total)))
        return fail(CCCS_EINVAL, "output buffers too small");
    size_t off = 0;
    for (size_t i = 0; i < all.size(); i++) {
        beams[i] = all[i].beam;
        if (uids) uids[i] = (int)all[i].uid;
        std::copy(all[i].w.begin(), all[i].w.end(), weights + off);
        off += all[i].w.size();
    }
    return CCCS_OK;
}

int cccs_h5_write_roi_list(const char* path, int nroi, const char* const* names, const uint8_t* const* masks,
                           const uint32_t size[3], const uint32_t* crop_start, const uint32_t* crop_size) {
    g_err.clear();
    if (!path || nroi < 0 || (nroi && (!names || !masks || !crop_start || !crop_size)) || !size)
        return fail(CCCS_EINVAL, "null argument");
    std::set<std::string> seen;
    for (int r = 0; r < nroi; r++) {
        if (!names[r] || !names[r][0] || std::strchr(names[r], '/') || !seen.insert(names[r]).second)
            return fail(CCCS_EINVAL, "ROI names must be unique, non-empty and free of '/'");
        for (int a = 0; a < 3; a++)
            if (size[a] > 65535 || crop_start[3 * r + a] > 65535 || crop_size[3 * r + a] > 65535)
                return fail(CCCS_EINVAL, "ROI sizes must fit 16 bits");
        if ((uint64_t)crop_size[3 * r] * crop_size[3 * r + 1] * crop_size[3 * r + 2] != 0 && !masks[r])
            return fail(CCCS_EINVAL, "null mask");
    }
    if (nroi > 65535) return fail(CCCS_EINVAL, "too many ROIs");
    std::lock_guard<std::mutex> lk(g_lock);
    quiet();
    Id f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(CCCS_EIO, std::string("cannot create ") + path);
    const unsigned vs[3] = {size[0], size[1], size[2]};
    for (int r = 0; r < nroi; r++) {
        Id g = group_create(f, names[r]);
        if (!g.ok()) return fail(CCCS_EIO, std::string("cannot create ROI group ") + names[r]);
        const unsigned cs[3] = {crop_size[3 * r], crop_size[3 * r + 1], crop_size[3 * r + 2]};
        const unsigned c0[3] = {crop_start[3 * r], crop_start[3 * r + 1], crop_start[3 * r + 2]};
        Id pg = group_create(g, "ArrayProps");
        const unsigned index = (unsigned)r + 1;
        bool ok = pg.ok() && attr_string(g, "name", names[r]) &&
                  attr_tuple3(pg, "size", H5T_STD_U16LE, H5T_NATIVE_UINT, vs) &&
                  attr_tuple3(pg, "crop_size", H5T_STD_U16LE, H5T_NATIVE_UINT, cs) &&
                  attr_tuple3(pg, "crop_start", H5T_STD_U16LE, H5T_NATIVE_UINT, c0);
        const hsize_t nv = (hsize_t)cs[0] * cs[1] * cs[2];
        if (ok && nv > 0) {
            Id s = simple_space(1, &nv);
            Id d(H5Dcreate2(g, "mask", H5T_STD_U8LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
            ok = d.ok() && H5Dwrite(d, H5T_NATIVE_UINT8, H5S_ALL, H5S_ALL, H5P_DEFAULT, masks[r]) >= 0;
        }
        ok = ok && attr_scalar(g, "index", H5T_STD_U16LE, H5T_NATIVE_UINT, &index);
        if (!ok) return fail(CCCS_EIO, std::string("cannot write ROI ") + names[r]);
    }
    if (H5Fflush(f, H5F_SCOPE_LOCAL) < 0) return fail(CCCS_EIO, "flush failed");
    return CCCS_OK;
}

int cccs_h5_read_roi_list(const char* path, int cap_roi, int* nroi, char (*names)[64], uint32_t size[3],
                          uint32_t* crop_start, uint32_t* crop_size, uint8_t* masks, size_t cap_bytes, size_t* nbytes) {
    g_err.clear();
    if (!path || !nroi || !nbytes || !size) return fail(CCCS_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(g_lock);
    quiet();
    Id f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(CCCS_EIO, std::string("cannot open ") + path);
    Id root = group_open(f, "/");
    std::vector<std::string> links;
    if (!list_links(root, links)) return fail(CCCS_EIO, "cannot list the ROI groups");
    struct Roi {
        unsigned index;
        std::string name;
        unsigned vs[3], c0[3], cs[3];
        std::vector<uint8_t> mask;
    };
    std::vector<Roi> rois;
    for (const std::string& ln : links) {
        Id g = group_open(root, ln.c_str());
        if (!g.ok()) return fail(CCCS_EIO, "cannot open ROI group " + ln);
        Roi r{};
        if (!attr_read(g, "index", H5T_NATIVE_UINT, &r.index) || !attr_read_string(g, "name", r.name))
            return fail(CCCS_EIO, "index / name missing in " + ln);
        Id pg = group_open(g, "ArrayProps");
        if (!pg.ok() || !attr_read_tuple3(pg, "size", H5T_NATIVE_UINT, r.vs) ||
            !attr_read_tuple3(pg, "crop_size", H5T_NATIVE_UINT, r.cs) ||
            !attr_read_tuple3(pg, "crop_start", H5T_NATIVE_UINT, r.c0))
            return fail(CCCS_EIO, "ArrayProps missing in " + ln);
        if (H5Lexists(g, "mask", H5P_DEFAULT) > 0) {
            Id d(H5Dopen2(g, "mask", H5P_DEFAULT), H5Dclose);
            Id s(H5Dget_space(d), H5Sclose);
            const hssize_t n = H5Sget_simple_extent_npoints(s);
            if (!d.ok() || n < 0) return fail(CCCS_EIO, "bad mask in " + ln);
            r.mask.resize((size_t)n);
            if (n && H5Dread(d, H5T_NATIVE_UINT8, H5S_ALL, H5S_ALL, H5P_DEFAULT, r.mask.data()) < 0)
                return fail(CCCS_EIO, "cannot read the mask of " + ln);
        }
        if ((uint64_t)r.cs[0] * r.cs[1] * r.cs[2] != r.mask.size())
            return fail(CCCS_EIO, "mask size disagrees with crop_size in " + ln);
        rois.push_back(std::move(r));
    }
    std::stable_sort(rois.begin(), rois.end(), [](const Roi& a, const Roi& b) { return a.index < b.index; });
    size_t total = 0;
    for (const Roi& r : rois) total += r.mask.size();
    *nroi = (int)rois.size();
    *nbytes = total;
    if (!rois.empty())
        for (int a = 0; a < 3; a++) size[a] = rois[0].vs[a];
    if (!masks) return CCCS_OK;
    if (cap_roi < (int)rois.size() || !names || !crop_start || !crop_size || cap_bytes < total)
        return fail(CCCS_EINVAL, "output buffers too small");
    size_t off = 0;
    for (size_t i = 0; i < rois.size(); i++) {
        std::snprintf(names[i], 64, "%s", rois[i].name.c_str());
        for (int a = 0; a < 3; a++) {
            crop_start[3 * i + a] = rois[i].c0[a];
            crop_size[3 * i + a] = rois[i].cs[a];
        }
        std::copy(rois[i].mask.begin(), rois[i].mask.end(), masks + off);
        off += rois[i].mask.size();
    }
    return CCCS_OK;
}

}  // extern "C"
