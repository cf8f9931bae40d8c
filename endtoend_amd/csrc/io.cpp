// io.cpp -- the reference's input/output file formats for the dose path: beam lists
// (DoseCalcIO/beam.cpp:369-607, reader and writer), the omni header
// (DoseCalcIO/calc_header.cpp:12-161), fluence maps (DoseCalcIO/fluence.cpp:19-50), raw float32
// volumes (density in, dose out), and the preprocessor's aperture-ready fill
// (dosecalc-preprocess/fmapProcessing.cpp:7-46).  Host code only.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/cccs.h"
#include "errors.hpp"

using namespace cccs;

namespace {

constexpr float kPiF = 3.141592654f;  // the reference's PI (include/dosecalc_defs.h:9)
constexpr float kDefaultSad = 100.f;  // DEFAULT_SAD (beam.cpp:26)

// Whitespace-free tokens separated by spaces (tokenize_string with " ", io_helpers.cpp:215-227:
// runs of spaces are one separator; tabs are not separators).
std::vector<std::string> split_spaces(const std::string& s) {
    std::vector<std::string> out;
    size_t p = s.find_first_not_of(' ');
    while (p != std::string::npos) {
        size_t q = s.find_first_of(' ', p);
        out.push_back(s.substr(p, q == std::string::npos ? std::string::npos : q - p));
        p = s.find_first_not_of(' ', q);
    }
    return out;
}

bool is_comment(const std::string& s) {  // io_helpers.cpp:207-213
    size_t f = s.find_first_not_of(" \t");
    return f != std::string::npos && s[f] == '#';
}

bool is_number(const std::string& s) {  // io_helpers.cpp:238-260: the whole token parses as a double
    if (s.empty()) return false;
    char* end = nullptr;
    errno = 0;
    std::strtod(s.c_str(), &end);
    return errno == 0 && end && *end == 0;
}

bool to_float(const std::string& s, float& v) {  // std::stof: leading number, else an error
    char* end = nullptr;
    errno = 0;
    v = std::strtof(s.c_str(), &end);
    return end != s.c_str() && errno == 0;
}

int reconfigure_or_fail(cccs_beam& b, int from_direction, int line) {
    int rc = cccs_beam_reconfigure(&b, from_direction);
    if (rc != CCCS_OK) return set_error(CCCS_EINVAL, "beam on line " + std::to_string(line) + ": " + cccs_last_error());
    return CCCS_OK;
}

}  // namespace

extern "C" int cccs_load_beam_list(const char* path, int requested, cccs_beam* beams, int cap, int* count) {
    clear_error();
    if (!path || !count || (cap > 0 && !beams) || cap < 0) return set_error(CCCS_EINVAL, "null argument");
    *count = 0;
    std::ifstream f(path);
    if (!f.is_open()) return set_error(CCCS_EIO, std::string("cannot open beam list ") + path);
    if (requested <= 0) {  // get_line_count (io_helpers.cpp)
        std::ifstream g(path);
        std::string l;
        requested = 0;
        while (std::getline(g, l)) requested++;
    }
    int n = 0;
    for (int line_no = 1; n < requested; line_no++) {
        std::string line;
        std::getline(f, line);
        // the reference stops when getline hits end-of-file, which also drops a last line
        // that has no trailing newline (beam.cpp:395-399)
        if (f.eof()) break;
        if (line.empty() || is_comment(line)) continue;
        std::vector<std::string> t = split_spaces(line);
        int N = (int)t.size() - 2, ptr = 0;
        if (N < 0) return set_error(CCCS_EINVAL, "beam list line " + std::to_string(line_no) + ": needs \"%f %f\"");
        cccs_beam b;
        std::memset(&b, 0, sizeof(b));
        b.sad = kDefaultSad;
        float az, ze;
        if (!to_float(t[ptr++], az) || !to_float(t[ptr++], ze))
            return set_error(CCCS_EINVAL, "beam list line " + std::to_string(line_no) + ": bad angle");
        b.azimuth = (az * kPiF) / 180.0f;
        b.zenith = (ze * kPiF) / 180.0f;
        int from_dir = 0;
        while (N > 0 && ptr < (int)t.size()) {
            if (t[ptr] == "iso:") {
                ptr++;
                if ((N -= 4) < 0) return set_error(CCCS_EINVAL, "line " + std::to_string(line_no) + ": iso: needs 3 values");
                for (int k = 0; k < 3; k++)
                    if (!to_float(t[ptr++], b.isocenter[k])) return set_error(CCCS_EINVAL, "bad iso value");
                continue;
            }
            if (t[ptr] == "sad:") {
                ptr++;
                if ((N -= 2) < 0) return set_error(CCCS_EINVAL, "line " + std::to_string(line_no) + ": sad: needs 1 value");
                if (!to_float(t[ptr++], b.sad)) return set_error(CCCS_EINVAL, "bad sad value");
                continue;
            }
            if (t[ptr] == "dir:") {
                ptr++;
                if ((N -= 4) < 0) return set_error(CCCS_EINVAL, "line " + std::to_string(line_no) + ": dir: needs 3 values");
                float d[3];
                for (int k = 0; k < 3; k++)
                    if (!to_float(t[ptr++], d[k])) return set_error(CCCS_EINVAL, "bad dir value");
                const float inv = 1.0f / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                for (int k = 0; k < 3; k++) b.direction[k] = d[k] * inv;
                from_dir = 1;
                continue;
            }
            if (is_number(t[ptr])) {  // collimator angle
                if ((N -= 1) < 0) return set_error(CCCS_EINVAL, "bad collimator field");
                float c;
                to_float(t[ptr++], c);
                b.coll = (c * kPiF) / 180.f;
                continue;
            }
            ++ptr;  // unknown field: skipped, as the reference does
            --N;
        }
        CCCS_TRY(reconfigure_or_fail(b, from_dir, line_no));
        if (n < cap) beams[n] = b;
        n++;
    }
    *count = n;
    if (n > cap) return set_error(CCCS_EINVAL, "beam list has more beams than the output array holds");
    return CCCS_OK;
}

extern "C" int cccs_load_omni_beam_list(const char* path, int beam_count, cccs_beam* beams, char* fmap_names,
                                        int name_len, int* count) {
    clear_error();
    if (!path || !count || beam_count < 0 || (beam_count > 0 && !beams)) return set_error(CCCS_EINVAL, "null argument");
    *count = 0;
    FILE* in = std::fopen(path, "r");
    if (!in) return set_error(CCCS_EIO, std::string("cannot open beam list ") + path);
    int n = 0;
    if (std::fscanf(in, "%d", &n) != 1) {
        std::fclose(in);
        return set_error(CCCS_EIO, "omni beam list: missing beam count");
    }
    if (beam_count < n) n = beam_count;  // beam.cpp:528
    for (int b = 0; b < n; b++) {
        int id, fx, fy;
        char iso_type[25], iso_loc[25], orient[25], fname[1024];
        float az, ze, co;
        cccs_beam& o = beams[b];
        std::memset(&o, 0, sizeof(o));
        const int got = std::fscanf(in, "\n%d %f %f %f %f %24s %24s %f %f %f %f %f %f %24s %f %f %f %f %f %d %d %1023s", &id,
                                    &az, &ze, &co, &o.sad, iso_type, iso_loc, &o.isocenter[0], &o.isocenter[1],
                                    &o.isocenter[2], &o.source[0], &o.source[1], &o.source[2], orient, &o.direction[0],
                                    &o.direction[1], &o.direction[2], &o.beamlet_size[0], &o.beamlet_size[1], &fx, &fy,
                                    fname);
        if (got != 22) {
            std::fclose(in);
            return set_error(CCCS_EIO, "omni beam list: malformed beam " + std::to_string(b));
        }
        o.azimuth = (az * kPiF) / 180.0f;
        o.zenith = (ze * kPiF) / 180.0f;
        o.coll = (co * kPiF) / 180.0f;
        o.fmap_size[0] = (uint32_t)fx;
        o.fmap_size[1] = (uint32_t)fy;
        if (fmap_names && name_len > 0) {
            std::strncpy(fmap_names + (size_t)b * name_len, fname, (size_t)name_len - 1);
            fmap_names[(size_t)b * name_len + name_len - 1] = 0;
        }
    }
    std::fclose(in);
    *count = n;
    return CCCS_OK;
}

extern "C" int cccs_load_omni_header(const char* path, cccs_problem* p, int* beam_count, char* beam_spec,
                                     char* target_structure, int* reduce) {
    clear_error();
    if (!path || !p) return set_error(CCCS_EINVAL, "null argument");
    FILE* in = std::fopen(path, "r");
    if (!in) return set_error(CCCS_EIO, std::string("cannot open omni header ") + path);
    char dump[150];
    while (true) {  // skip the comment block (calc_header.cpp:92-97)
        if (!std::fgets(dump, sizeof(dump), in)) {
            std::fclose(in);
            return set_error(CCCS_EIO, "omni header: no data lines");
        }
        if (*dump != '#' && *dump != ' ') break;
    }
    int nphi = 0, ntheta = 0, nradii = 0, bc = 0, red = 0;
    unsigned ss = 0;
    char spec[100] = {0}, target[100] = {0};
    bool ok = std::fscanf(in, "%u %u %u", &p->size[0], &p->size[1], &p->size[2]) == 3;
    ok = ok && std::fscanf(in, "%f %f %f", &p->voxel[0], &p->voxel[1], &p->voxel[2]) == 3;
    ok = ok && std::fscanf(in, "%f %f %f", &p->start[0], &p->start[1], &p->start[2]) == 3;
    ok = ok && std::fscanf(in, "%u %u %u", &p->bbox_start[0], &p->bbox_start[1], &p->bbox_start[2]) == 3;
    ok = ok && std::fscanf(in, "%u %u %u", &p->bbox_size[0], &p->bbox_size[1], &p->bbox_size[2]) == 3;
    ok = ok && std::fscanf(in, "%u %u %u", &p->max_rev_size[0], &p->max_rev_size[1], &p->max_rev_size[2]) == 3;
    ok = ok && std::fscanf(in, "%f", &p->rev_latspacing) == 1;
    ok = ok && std::fscanf(in, "%f", &p->rev_longspacing) == 1;
    ok = ok && std::fscanf(in, "%f", &p->kernel_extent) == 1;
    ok = ok && std::fscanf(in, "%u", &ss) == 1;
    ok = ok && std::fscanf(in, "%d %d %d", &nphi, &ntheta, &nradii) == 3;
    ok = ok && std::fscanf(in, "%f", &p->penumbra) == 1;
    ok = ok && std::fscanf(in, "%d", &bc) == 1;
    ok = ok && std::fgets(dump, sizeof(dump), in) && std::fgets(spec, 99, in) && std::fgets(target, 99, in);
    ok = ok && std::fscanf(in, "%d", &red) == 1;
    std::fclose(in);
    if (!ok) return set_error(CCCS_EIO, "omni header: malformed");
    // strip the newline fgets keeps (calc_header.cpp:119-122)
    if (std::strlen(spec)) spec[std::strlen(spec) - 1] = 0;
    if (std::strlen(target)) target[std::strlen(target) - 1] = 0;
    p->ss_factor = ss;
    p->nphi = (uint32_t)nphi;
    p->ntheta = (uint32_t)ntheta;
    p->nradii = (uint32_t)nradii;
    // clamp the calc_bbox into the volume (calc_header.cpp:130-135)
    for (int k = 0; k < 3; k++) {
        if (p->bbox_start[k] >= p->size[k]) p->bbox_start[k] = 0;
        if (p->bbox_size[k] + p->bbox_start[k] > p->size[k]) p->bbox_size[k] = p->size[k] - p->bbox_start[k];
    }
    if (beam_count) *beam_count = bc;
    if (beam_spec) std::strcpy(beam_spec, spec);
    if (target_structure) std::strcpy(target_structure, target);
    if (reduce) *reduce = red;
    return CCCS_OK;
}

extern "C" int cccs_write_omni_header(const char* path, const cccs_problem* p, int beam_count, const char* beam_spec,
                                      const char* target_structure, int reduce) {
    clear_error();
    if (!path || !p || !beam_spec || !target_structure) return set_error(CCCS_EINVAL, "null argument");
    FILE* o = std::fopen(path, "w");
    if (!o) return set_error(CCCS_EIO, std::string("cannot write ") + path);
    // the field order and formats of write_omni_header (calc_header.cpp:12-80)
    std::fprintf(o, "# OMNI-HEADER: Nx Ny Nz | dx dy dz | x y z | bbox start | bbox size | REV dims | convlat | "
                    "convstep | kernel_extent | ss_factor | nphi ntheta nradii | penumbra | beam_count | spectrum | "
                    "target | reduce\n\n");
    std::fprintf(o, "%u %u %u\n", p->size[0], p->size[1], p->size[2]);
    std::fprintf(o, "%f %f %f\n", p->voxel[0], p->voxel[1], p->voxel[2]);
    std::fprintf(o, "%f %f %f\n", p->start[0], p->start[1], p->start[2]);
    std::fprintf(o, "%u %u %u\n", p->bbox_start[0], p->bbox_start[1], p->bbox_start[2]);
    std::fprintf(o, "%u %u %u\n", p->bbox_size[0], p->bbox_size[1], p->bbox_size[2]);
    std::fprintf(o, "%u %u %u\n", p->max_rev_size[0], p->max_rev_size[1], p->max_rev_size[2]);
    std::fprintf(o, "%f\n", p->rev_latspacing);
    std::fprintf(o, "%f\n", p->rev_longspacing);
    std::fprintf(o, "%f\n", p->kernel_extent);
    std::fprintf(o, "%u\n", p->ss_factor);
    std::fprintf(o, "%d %d %d\n", (int)p->nphi, (int)p->ntheta, (int)p->nradii);
    std::fprintf(o, "%f\n", p->penumbra);
    std::fprintf(o, "%d\n", beam_count);
    std::fprintf(o, "%s\n", beam_spec);
    std::fprintf(o, "%s\n", target_structure);
    std::fprintf(o, "%d\n", reduce);
    const bool bad = std::ferror(o) != 0;
    std::fclose(o);
    return bad ? set_error(CCCS_EIO, "write failed") : CCCS_OK;
}

extern "C" int cccs_read_f32(const char* path, float* out, uint64_t n) {
    clear_error();
    if (!path || (!out && n > 0)) return set_error(CCCS_EINVAL, "null argument");
    FILE* in = std::fopen(path, "rb");
    if (!in) return set_error(CCCS_EIO, std::string("cannot open ") + path);
    const size_t got = std::fread(out, sizeof(float), (size_t)n, in);
    std::fclose(in);
    // the fluence loader rejects a file of the wrong size (fluence.cpp:36-40)
    if (got != (size_t)n) return set_error(CCCS_EIO, "file has " + std::to_string(got) + " of " + std::to_string(n) + " floats");
    return CCCS_OK;
}

extern "C" int cccs_write_f32(const char* path, const float* data, uint64_t n) {
    clear_error();
    if (!path || (!data && n > 0)) return set_error(CCCS_EINVAL, "null argument");
    FILE* o = std::fopen(path, "wb");
    if (!o) return set_error(CCCS_EIO, std::string("cannot write ") + path);
    const size_t put = std::fwrite(data, sizeof(float), (size_t)n, o);
    std::fclose(o);
    return put == (size_t)n ? CCCS_OK : set_error(CCCS_EIO, "short write");
}

extern "C" int cccs_write_omni_beam_list(const char* path, const cccs_beam* beams, int beam_count,
                                         const char* const* iso_type, const char* const* iso_loc,
                                         const char* const* orient_type, const char* const* fmap_names) {
    clear_error();
    if (!path || (!beams && beam_count > 0) || beam_count < 0) return set_error(CCCS_EINVAL, "null argument");
    FILE* o = std::fopen(path, "w");
    if (!o) return set_error(CCCS_EIO, std::string("cannot write ") + path);
    std::fprintf(o, "%d\n", beam_count);
    char name[32];
    for (int b = 0; b < beam_count; b++) {
        const cccs_beam& B = beams[b];
        // deg_azimuth() and friends: 180 * angle / PI in FP32 (beam.h:52-54)
        const float az = 180.0f * B.azimuth / kPiF, zen = 180.0f * B.zenith / kPiF, coll = 180.0f * B.coll / kPiF;
        std::snprintf(name, sizeof(name), "fmap-%06d.raw", b);  // write_fluence_map's name (fluence.cpp:46-50)
        std::fprintf(o, "%d %f %f %f %f %s %s %f %f %f %f %f %f %s %f %f %f %f %f %d %d %s\n", b, az, zen, coll, B.sad,
                     iso_type && iso_type[b] ? iso_type[b] : "unspec", iso_loc && iso_loc[b] ? iso_loc[b] : "unspec",
                     B.isocenter[0], B.isocenter[1], B.isocenter[2], B.source[0], B.source[1], B.source[2],
                     orient_type && orient_type[b] ? orient_type[b] : "auto", B.direction[0], B.direction[1],
                     B.direction[2], B.beamlet_size[0], B.beamlet_size[1], (int)B.fmap_size[0], (int)B.fmap_size[1],
                     fmap_names && fmap_names[b] ? fmap_names[b] : name);
    }
    std::fprintf(o, "\n#beam_id  azimuth zenith coll sad | iso_type iso_loc isocenter.x isocenter.y isocenter.z | "
                    "source.x source.y source.z | orient_type direction.x direction.y direction.z | beamlet_size.x "
                    "beamlet_size.y fluence_dim.x fluence_dim.y | fmap_fname\n");
    const bool bad = std::ferror(o) != 0;
    std::fclose(o);
    return bad ? set_error(CCCS_EIO, "write failed") : CCCS_OK;
}

namespace {
// get_active_segments (dosecalc-preprocess/fmapProcessing.cpp:7-24): a non-zero beamlet restarts
// the run at itself, so each run is recorded as {last, last} of its non-zero beamlets.
struct Run {
    int first, last;
};
void row_runs(const float* f, int w, int yy, std::vector<Run>& out) {
    out.clear();
    int begin = -1, end = -1;
    for (int xx = 0; xx < w; ++xx) {
        const float v = f[(size_t)w * yy + xx];
        if (v != 0.f) begin = end = xx;
        if (begin >= 0) {
            if (v != 0.f) end = xx;
            if (v == 0.f || xx >= w - 1) {
                out.push_back(Run{begin, end});
                begin = end = -1;
            }
        }
    }
}
}  // namespace

extern "C" int cccs_fmap_aperture_fill(float* fluence, int width, int height, float fill_val) {
    clear_error();
    if (!fluence || width < 0 || height < 0) return set_error(CCCS_EINVAL, "bad argument");
    std::vector<Run> runs;
    for (int yy = 0; yy < height; ++yy) {
        row_runs(fluence, width, yy, runs);
        for (int ii = 0; ii + 1 < (int)runs.size(); ++ii)
            for (int xx = runs[ii].last + 1; xx < runs[ii + 1].first; ++xx) fluence[(size_t)width * yy + xx] = fill_val;
    }
    return CCCS_OK;
}

extern "C" int cccs_fmap_is_aperture_ready(const float* fluence, int width, int height, int* ready) {
    clear_error();
    if (!fluence || !ready || width < 0 || height < 0) return set_error(CCCS_EINVAL, "bad argument");
    std::vector<Run> runs;
    *ready = 1;
    for (int yy = 0; yy < height && *ready; ++yy) {
        row_runs(fluence, width, yy, runs);
        if (runs.size() > 1) *ready = 0;
    }
    return CCCS_OK;
}
