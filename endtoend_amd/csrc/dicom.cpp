// dicom.cpp -- DICOM ingestion on the host: CT/MR image series and RTSTRUCT contours read from a
// directory of DICOM Part 10 files, the structure helpers of the preprocess step (ROI lookup,
// extents, centroid) and the per-slice contour rasterisation.  Replaces the DCMTK-based readers
// RTImage (CCCS/RTClasses/rtimages.cpp:20-400) and RTStruct (rtstruct.cpp:100-380) and the host
// parts of dosecalc-preprocess (main.cu:139-320, DoseCalcIO/roi.cpp:201-345,
// cudaMemFunctions.cu:218-255).  Host code only; the resampling kernels are in ingest.hip.
//
// The reader follows the published encoding (DICOM PS3.5 section 7, PS3.10 section 7): a 128-byte
// preamble and "DICM", the file meta group in explicit VR little endian, then the data set in the
// transfer syntax it names (explicit or implicit VR little endian; other syntaxes are rejected).
// Sequences and items of defined or undefined length are parsed into a tree; elements are kept as
// raw bytes and decoded on access (US, SS, FD, FL binary; IS, DS, UI, CS, LO strings).
#include <dirent.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cccs.h"
#include "errors.hpp"

using namespace cccs;

namespace {

constexpr uint32_t tag(uint16_t g, uint16_t e) { return (uint32_t)g << 16 | e; }
constexpr uint32_t kItem = tag(0xFFFE, 0xE000), kItemEnd = tag(0xFFFE, 0xE00D), kSeqEnd = tag(0xFFFE, 0xE0DD);
constexpr uint32_t kUndefined = 0xFFFFFFFFu;

const char* const kCtSop = "1.2.840.10008.5.1.4.1.1.2";
const char* const kMrSop = "1.2.840.10008.5.1.4.1.1.4";
const char* const kRtImageSop = "1.2.840.10008.5.1.4.1.1.481.1";
const char* const kRtStructSop = "1.2.840.10008.5.1.4.1.1.481.3";

struct DataSet;
struct Element {
    char vr[2] = {0, 0};
    std::vector<uint8_t> value;
    std::vector<DataSet> items;  // sequences
};
struct DataSet {
    std::map<uint32_t, Element> el;
    const Element* find(uint32_t t) const {
        auto it = el.find(t);
        return it == el.end() ? nullptr : &it->second;
    }
};

// tags whose value is a sequence (needed to parse implicit VR data sets)
bool is_sequence_tag(uint32_t t) {
    switch (t) {
        case tag(0x3006, 0x0010): case tag(0x3006, 0x0012): case tag(0x3006, 0x0014): case tag(0x3006, 0x0016):
        case tag(0x3006, 0x0020): case tag(0x3006, 0x0039): case tag(0x3006, 0x0040): case tag(0x3006, 0x0080):
            return true;
        default: return false;
    }
}

bool long_length_vr(const char* vr) {
    static const char* const kLong[] = {"OB", "OD", "OF", "OL", "OW", "SQ", "UC", "UR", "UT", "UN", "OV", "SV", "UV"};
    for (const char* l : kLong)
        if (vr[0] == l[0] && vr[1] == l[1]) return true;
    return false;
}

struct Reader {
    const uint8_t* p;
    size_t n, pos = 0;
    bool explicit_vr = true;
    std::string err;
    bool need(size_t k) {
        if (pos + k > n) {
            err = "truncated element";
            return false;
        }
        return true;
    }
    uint16_t u16() { uint16_t v; std::memcpy(&v, p + pos, 2); pos += 2; return v; }
    uint32_t u32() { uint32_t v; std::memcpy(&v, p + pos, 4); pos += 4; return v; }

    // one data set up to `end` (a byte offset) or, for undefined length, up to the delimiter `stop`
    bool dataset(DataSet& ds, size_t end, uint32_t stop, int depth) {
        if (depth > 16) {
            err = "sequences nested too deep";
            return false;
        }
        while (pos < end) {
            if (!need(8)) return false;
            const uint16_t g = u16(), e = u16();
            const uint32_t t = (uint32_t)g << 16 | e;
            if (stop && t == stop) {
                u32();  // zero length of the delimiter
                return true;
            }
            Element el;
            uint32_t len;
            // item tags carry no VR in any syntax, and group 0002 is always explicit
            const bool expl = g == 0x0002 || (explicit_vr && g != 0xFFFE);
            if (expl) {
                el.vr[0] = (char)p[pos];
                el.vr[1] = (char)p[pos + 1];
                pos += 2;
                if (long_length_vr(el.vr)) {
                    if (!need(6)) return false;
                    pos += 2;
                    len = u32();
                } else {
                    if (!need(2)) return false;
                    len = u16();
                }
            } else {
                len = u32();
                if (is_sequence_tag(t) || len == kUndefined) {
                    el.vr[0] = 'S';
                    el.vr[1] = 'Q';
                }
            }
            const bool sq = el.vr[0] == 'S' && el.vr[1] == 'Q';
            if (sq || (len == kUndefined && t != tag(0x7FE0, 0x0010))) {
                if (!sequence(el, len, depth)) return false;
            } else if (len == kUndefined) {
                err = "encapsulated (compressed) pixel data is not supported";
                return false;
            } else {
                if (!need(len)) return false;
                el.value.assign(p + pos, p + pos + len);
                pos += len;
            }
            ds.el[t] = std::move(el);
        }
        return true;
    }

    bool sequence(Element& el, uint32_t len, int depth) {
        const size_t end = len == kUndefined ? n : pos + len;
        if (len != kUndefined && end > n) {
            err = "sequence beyond the end of the file";
            return false;
        }
        while (pos < end) {
            if (!need(8)) return false;
            const uint16_t g = u16(), e = u16();
            const uint32_t t = (uint32_t)g << 16 | e;
            const uint32_t ilen = u32();
            if (t == kSeqEnd) return true;
            if (t != kItem) {
                err = "expected a sequence item";
                return false;
            }
            DataSet item;
            if (ilen == kUndefined) {
                if (!dataset(item, n, kItemEnd, depth + 1)) return false;
            } else {
                const size_t iend = pos + ilen;
                if (iend > end) {
                    err = "item beyond its sequence";
                    return false;
                }
                if (!dataset(item, iend, 0, depth + 1)) return false;
                pos = iend;
            }
            el.items.push_back(std::move(item));
        }
        return true;
    }
};

std::string trim(std::string s) {
    while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
    size_t k = 0;
    while (k < s.size() && s[k] == ' ') k++;
    return s.substr(k);
}

std::string as_string(const Element* e) {
    return e ? trim(std::string(e->value.begin(), e->value.end())) : std::string();
}

// backslash-separated decimal strings (DS, IS) or binary FD/FL values
std::vector<double> as_numbers(const Element* e) {
    std::vector<double> v;
    if (!e) return v;
    if (e->vr[0] == 'F' && e->vr[1] == 'D') {
        for (size_t k = 0; k + 8 <= e->value.size(); k += 8) {
            double d;
            std::memcpy(&d, e->value.data() + k, 8);
            v.push_back(d);
        }
        return v;
    }
    if (e->vr[0] == 'F' && e->vr[1] == 'L') {
        for (size_t k = 0; k + 4 <= e->value.size(); k += 4) {
            float f;
            std::memcpy(&f, e->value.data() + k, 4);
            v.push_back(f);
        }
        return v;
    }
    const std::string s(e->value.begin(), e->value.end());
    size_t a = 0;
    while (a <= s.size()) {
        size_t b = s.find('\\', a);
        if (b == std::string::npos) b = s.size();
        const std::string f = trim(s.substr(a, b - a));
        if (!f.empty()) v.push_back(std::strtod(f.c_str(), nullptr));
        a = b + 1;
    }
    return v;
}

bool as_uint16(const Element* e, uint16_t& out) {
    if (!e) return false;
    if ((e->vr[0] == 'U' && e->vr[1] == 'S') || (e->vr[0] == 0 && e->value.size() == 2)) {
        if (e->value.size() < 2) return false;
        std::memcpy(&out, e->value.data(), 2);
        return true;
    }
    const std::vector<double> v = as_numbers(e);
    if (v.empty()) return false;
    out = (uint16_t)v[0];
    return true;
}

struct Parsed {
    DataSet ds;
    std::string transfer;
};

int parse_file(const std::string& path, Parsed& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return set_error(CCCS_EIO, "cannot open " + path);
    std::vector<uint8_t> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    Reader r{buf.data(), buf.size()};
    if (buf.size() >= 132 && std::memcmp(buf.data() + 128, "DICM", 4) == 0) {
        r.pos = 132;
        // file meta group: explicit VR little endian, up to the first element of another group
        while (r.pos + 8 <= r.n) {
            uint16_t g;
            std::memcpy(&g, buf.data() + r.pos, 2);
            if (g != 0x0002) break;
            Reader m{buf.data(), buf.size()};
            m.pos = r.pos;
            const uint16_t gg = m.u16(), ee = m.u16();
            char vr[2] = {(char)buf[m.pos], (char)buf[m.pos + 1]};
            m.pos += 2;
            uint32_t len;
            if (long_length_vr(vr)) {
                m.pos += 2;
                len = m.u32();
            } else {
                len = m.u16();
            }
            if (m.pos + len > m.n) return set_error(CCCS_EIO, path + ": truncated file meta group");
            if (gg == 0x0002 && ee == 0x0010) out.transfer = trim(std::string(buf.begin() + m.pos, buf.begin() + m.pos + len));
            r.pos = m.pos + len;
        }
    } else {
        out.transfer = "1.2.840.10008.1.2";  // no preamble: a bare implicit VR little endian data set
    }
    if (out.transfer == "1.2.840.10008.1.2") r.explicit_vr = false;
    else if (out.transfer == "1.2.840.10008.1.2.1" || out.transfer.empty()) r.explicit_vr = true;
    else return set_error(CCCS_EINVAL, path + ": unsupported transfer syntax " + out.transfer);
    if (!r.dataset(out.ds, r.n, 0, 0)) return set_error(CCCS_EIO, path + ": " + r.err);
    return CCCS_OK;
}

std::vector<std::string> list_dir(const std::string& dir, bool dcm_only) {
    std::vector<std::string> names;
    DIR* d = opendir(dir.c_str());
    if (!d) return names;
    while (dirent* e = readdir(d)) {
        if (!std::strcmp(e->d_name, ".") || !std::strcmp(e->d_name, "..")) continue;
        if (dcm_only && !std::strstr(e->d_name, ".dcm")) continue;
        names.push_back(e->d_name);
    }
    closedir(d);
    std::sort(names.begin(), names.end());  // deterministic order (readdir's is unspecified)
    return names;
}

struct Slice {
    std::string file;
    float ipp[3];
};

// The image files of `dir` (names containing ".dcm" with a CT, MR or RT image SOP class), sorted
// from foot to head by ImagePositionPatient z (RTImage::loadDicomInfo, rtimages.cpp:20-84).
int scan_series(const std::string& dir, std::vector<Slice>& slices, int& sop) {
    slices.clear();
    sop = 0;
    DIR* d = opendir(dir.c_str());
    if (!d) return set_error(CCCS_EIO, "cannot open " + dir);
    closedir(d);
    for (const std::string& n : list_dir(dir, true)) {
        Parsed ps;
        if (parse_file(dir + "/" + n, ps) != CCCS_OK) {
            clear_error();
            continue;  // the reference warns and skips unreadable files
        }
        const std::string cls = as_string(ps.ds.find(tag(0x0008, 0x0016)));
        int kind = cls == kCtSop ? 1 : cls == kMrSop ? 2 : cls == kRtImageSop ? 3 : 0;
        if (!kind) continue;
        sop = kind;
        Slice s;
        s.file = dir + "/" + n;
        const std::vector<double> ipp = as_numbers(ps.ds.find(tag(0x0020, 0x0032)));
        for (int k = 0; k < 3; k++) s.ipp[k] = k < (int)ipp.size() ? (float)ipp[k] : 0.f;
        slices.push_back(s);
    }
    if (slices.empty()) return set_error(CCCS_EINVAL, "no CT/MR/RT image files (*.dcm) in " + dir);
    std::stable_sort(slices.begin(), slices.end(), [](const Slice& a, const Slice& b) { return a.ipp[2] < b.ipp[2]; });
    return CCCS_OK;
}

// polygon fill on one slice: edges drawn as 8-connected lines between the integer vertices, the
// interior by scanlines (rows y with an edge spanning [y0, y1) contribute the crossing at
// x0 + (y - y0)(x1 - x0)/(y1 - y0); pixels from ceil(left) to floor(right) of each pair are set)
void draw_line(uint8_t* img, int w, int h, int x0, int y0, int x1, int y1) {
    const int dx = std::abs(x1 - x0), dy = std::abs(y1 - y0), sx = x0 < x1 ? 1 : -1, sy = y0 < y1 ? 1 : -1;
    int err = dx - dy;
    for (;;) {
        if (x0 >= 0 && x0 < w && y0 >= 0 && y0 < h) img[(size_t)y0 * w + x0] = 1;
        if (x0 == x1 && y0 == y1) break;
        const int e2 = 2 * err;
        if (e2 > -dy) { err -= dy; x0 += sx; }
        if (e2 < dx) { err += dx; y0 += sy; }
    }
}

void fill_polygon(uint8_t* img, int w, int h, const std::vector<int>& xs, const std::vector<int>& ys) {
    const int n = (int)xs.size();
    if (n == 0) return;
    for (int i = 0; i < n; i++) {
        const int j = (i + 1) % n;
        draw_line(img, w, h, xs[i], ys[i], xs[j], ys[j]);
    }
    if (n < 3) return;
    const int ymin = std::max(0, *std::min_element(ys.begin(), ys.end()));
    const int ymax = std::min(h - 1, *std::max_element(ys.begin(), ys.end()));
    std::vector<std::pair<int64_t, int64_t>> cross;  // crossing x as a fraction num/den (den > 0)
    for (int y = ymin; y <= ymax; y++) {
        cross.clear();
        for (int i = 0; i < n; i++) {
            const int j = (i + 1) % n;
            int x0 = xs[i], y0 = ys[i], x1 = xs[j], y1 = ys[j];
            if (y0 == y1) continue;
            if (y0 > y1) { std::swap(x0, x1); std::swap(y0, y1); }
            if (y < y0 || y >= y1) continue;
            const int64_t den = y1 - y0;
            cross.emplace_back((int64_t)x0 * den + (int64_t)(y - y0) * (x1 - x0), den);
        }
        std::sort(cross.begin(), cross.end(), [](const std::pair<int64_t, int64_t>& a, const std::pair<int64_t, int64_t>& b) {
            return (__int128)a.first * b.second < (__int128)b.first * a.second;
        });
        for (size_t k = 0; k + 1 < cross.size(); k += 2) {
            auto fl = [](int64_t a, int64_t b) { return a >= 0 ? a / b : -((-a + b - 1) / b); };
            const int64_t xa = -fl(-cross[k].first, cross[k].second);      // ceil
            const int64_t xb = fl(cross[k + 1].first, cross[k + 1].second);  // floor
            for (int64_t x = std::max<int64_t>(xa, 0); x <= std::min<int64_t>(xb, w - 1); x++) img[(size_t)y * w + x] = 1;
        }
    }
}

}  // namespace

struct cccs_rtstruct {
    struct Roi {
        std::string name;
        int number = 0;
        std::vector<uint32_t> counts;
        std::vector<float> points;  // mm, x y z per point, sub-contours back to back
    };
    std::vector<Roi> rois;
};

extern "C" {

int cccs_dicom_ct_info(const char* dir, cccs_ct_info* out) {
    clear_error();
    if (!dir || !out) return set_error(CCCS_EINVAL, "null argument");
    std::memset(out, 0, sizeof(*out));
    std::vector<Slice> slices;
    int sop = 0;
    CCCS_TRY(scan_series(dir, slices, sop));
    Parsed ps;
    CCCS_TRY(parse_file(slices[0].file, ps));
    const DataSet& ds = ps.ds;
    uint16_t cols = 0, rows = 0;
    if (!as_uint16(ds.find(tag(0x0028, 0x0011)), cols) || !as_uint16(ds.find(tag(0x0028, 0x0010)), rows))
        return set_error(CCCS_EINVAL, "image without Rows/Columns");
    out->size[0] = cols;
    out->size[1] = rows;
    out->size[2] = (uint32_t)slices.size();
    // the reference takes both in-plane spacings from PixelSpacing[0] (rtimages.cpp:273-276)
    const std::vector<double> ps_mm = as_numbers(ds.find(tag(0x0028, 0x0030)));
    if (!ps_mm.empty()) out->voxel_mm[0] = out->voxel_mm[1] = (float)ps_mm[0];
    // slice spacing from the first two positions, else SliceThickness (:278-285)
    if (slices.size() >= 2) out->voxel_mm[2] = slices[1].ipp[2] - slices[0].ipp[2];
    else {
        const std::vector<double> st = as_numbers(ds.find(tag(0x0018, 0x0050)));
        if (!st.empty()) out->voxel_mm[2] = (float)st[0];
    }
    for (int k = 0; k < 3; k++) out->start_mm[k] = slices[0].ipp[k];
    const std::vector<double> io = as_numbers(ds.find(tag(0x0020, 0x0037)));
    for (int k = 0; k < 6 && k < (int)io.size(); k++) out->orientation[k] = (float)io[k];
    out->rescale_slope = 1.f;
    out->rescale_intercept = 0.f;
    if (sop == 1) {
        const std::vector<double> sl = as_numbers(ds.find(tag(0x0028, 0x1053)));
        const std::vector<double> ic = as_numbers(ds.find(tag(0x0028, 0x1052)));
        if (!sl.empty()) out->rescale_slope = (float)sl[0];
        if (!ic.empty()) out->rescale_intercept = (float)ic[0];
    }
    out->sop_type = sop;
    return CCCS_OK;
}

int cccs_dicom_ct_read(const char* dir, float* hu, uint64_t cap) {
    clear_error();
    if (!dir || !hu) return set_error(CCCS_EINVAL, "null argument");
    cccs_ct_info info;
    CCCS_TRY(cccs_dicom_ct_info(dir, &info));
    const uint64_t plane = (uint64_t)info.size[0] * info.size[1];
    if (cap < plane * info.size[2]) return set_error(CCCS_EINVAL, "output buffer too small");
    std::vector<Slice> slices;
    int sop = 0;
    CCCS_TRY(scan_series(dir, slices, sop));
    for (size_t i = 0; i < slices.size(); i++) {
        Parsed ps;
        CCCS_TRY(parse_file(slices[i].file, ps));
        const Element* px = ps.ds.find(tag(0x7FE0, 0x0010));
        uint16_t bits = 16, repr = 0, r = 0, c = 0;
        as_uint16(ps.ds.find(tag(0x0028, 0x0100)), bits);
        as_uint16(ps.ds.find(tag(0x0028, 0x0103)), repr);
        as_uint16(ps.ds.find(tag(0x0028, 0x0010)), r);
        as_uint16(ps.ds.find(tag(0x0028, 0x0011)), c);
        if (!px || px->value.size() < plane * 2) return set_error(CCCS_EINVAL, slices[i].file + ": missing pixel data");
        if (bits != 16) return set_error(CCCS_EINVAL, slices[i].file + ": only 16-bit pixels are supported");
        if (r != info.size[1] || c != info.size[0]) return set_error(CCCS_EINVAL, slices[i].file + ": slice size differs");
        // value * slope + intercept in FP32, two's complement when PixelRepresentation is 1
        // (rtimages.cpp:370-390)
        const uint8_t* src = px->value.data();
        float* dst = hu + plane * i;
        for (uint64_t k = 0; k < plane; k++) {
            uint16_t u;
            std::memcpy(&u, src + 2 * k, 2);
            float v = repr == 1 ? (float)(int16_t)u : (float)u;
            if (info.sop_type == 1) {
                v *= info.rescale_slope;
                v += info.rescale_intercept;
            }
            dst[k] = v;
        }
    }
    return CCCS_OK;
}

int cccs_dicom_rtstruct_open(const char* dir, cccs_rtstruct** out) {
    clear_error();
    if (!dir || !out) return set_error(CCCS_EINVAL, "null argument");
    *out = nullptr;
    // the first file with the RT Structure Set SOP class (any name, RTStruct::loadDicomInfo)
    Parsed ps;
    bool found = false;
    for (const std::string& n : list_dir(dir, false)) {
        Parsed p;
        if (parse_file(std::string(dir) + "/" + n, p) != CCCS_OK) {
            clear_error();
            continue;
        }
        if (as_string(p.ds.find(tag(0x0008, 0x0016))) == kRtStructSop) {
            ps = std::move(p);
            found = true;
            break;
        }
    }
    if (!found) return set_error(CCCS_EINVAL, std::string("no RTSTRUCT file in ") + dir);
    std::unique_ptr<cccs_rtstruct> s(new cccs_rtstruct());
    const Element* roiseq = ps.ds.find(tag(0x3006, 0x0020));
    const Element* cseq = ps.ds.find(tag(0x3006, 0x0039));
    const size_t nroi = roiseq ? roiseq->items.size() : 0;
    s->rois.resize(nroi);
    for (size_t r = 0; r < nroi; r++) {
        cccs_rtstruct::Roi& roi = s->rois[r];
        const DataSet& it = roiseq->items[r];
        roi.name = as_string(it.find(tag(0x3006, 0x0026)));
        const std::vector<double> num = as_numbers(it.find(tag(0x3006, 0x0022)));
        roi.number = num.empty() ? 0 : (int)num[0];
        // the contours of ROI r are item r of the ROI Contour Sequence (rtstruct.cpp:268, 340)
        if (!cseq || r >= cseq->items.size()) continue;
        const Element* contours = cseq->items[r].find(tag(0x3006, 0x0040));
        if (!contours) continue;
        for (const DataSet& c : contours->items) {
            const std::vector<double> np = as_numbers(c.find(tag(0x3006, 0x0046)));
            const std::vector<double> pts = as_numbers(c.find(tag(0x3006, 0x0050)));
            const uint32_t count = np.empty() ? (uint32_t)(pts.size() / 3) : (uint32_t)np[0];
            if ((size_t)count * 3 > pts.size()) return set_error(CCCS_EINVAL, "ContourData shorter than NumberOfContourPoints");
            roi.counts.push_back(count);
            for (uint32_t k = 0; k < 3 * count; k++) roi.points.push_back((float)pts[k]);
        }
    }
    *out = s.release();
    return CCCS_OK;
}

int cccs_rtstruct_count(const cccs_rtstruct* s, int* nroi) {
    clear_error();
    if (!s || !nroi) return set_error(CCCS_EINVAL, "null argument");
    *nroi = (int)s->rois.size();
    return CCCS_OK;
}

int cccs_rtstruct_roi_info(const cccs_rtstruct* s, int roi, char* name, int name_cap, int* roi_number, int* nsub,
                           uint64_t* npoints) {
    clear_error();
    if (!s) return set_error(CCCS_EINVAL, "null argument");
    if (roi < 0 || roi >= (int)s->rois.size()) return set_error(CCCS_EINVAL, "ROI index out of range");
    const cccs_rtstruct::Roi& r = s->rois[roi];
    if (name && name_cap > 0) {
        std::snprintf(name, (size_t)name_cap, "%s", r.name.c_str());
    }
    if (roi_number) *roi_number = r.number;
    if (nsub) *nsub = (int)r.counts.size();
    if (npoints) *npoints = r.points.size() / 3;
    return CCCS_OK;
}

int cccs_rtstruct_roi_points(const cccs_rtstruct* s, int roi, uint32_t* sub_counts, float* points_mm) {
    clear_error();
    if (!s) return set_error(CCCS_EINVAL, "null argument");
    if (roi < 0 || roi >= (int)s->rois.size()) return set_error(CCCS_EINVAL, "ROI index out of range");
    const cccs_rtstruct::Roi& r = s->rois[roi];
    if (sub_counts && !r.counts.empty()) std::memcpy(sub_counts, r.counts.data(), r.counts.size() * sizeof(uint32_t));
    if (points_mm && !r.points.empty()) std::memcpy(points_mm, r.points.data(), r.points.size() * sizeof(float));
    return CCCS_OK;
}

int cccs_rtstruct_find(const cccs_rtstruct* s, const char* search, int exact, int* roi) {
    clear_error();
    if (!s || !search || !roi) return set_error(CCCS_EINVAL, "null argument");
    // getROIIndex (DoseCalcIO/roi.cpp:201-215): first exact or substring match, -1 when none
    *roi = -1;
    for (size_t r = 0; r < s->rois.size(); r++) {
        const std::string& n = s->rois[r].name;
        if (exact ? n == search : n.find(search) != std::string::npos) {
            *roi = (int)r;
            break;
        }
    }
    return CCCS_OK;
}

int cccs_rtstruct_close(cccs_rtstruct* s) {
    delete s;
    return CCCS_OK;
}

int cccs_roi_extents(const float* points_mm, uint64_t npoints, const uint32_t frame_size[3],
                     const float frame_start_cm[3], const float frame_spacing_cm[3], uint32_t crop_start[3],
                     uint32_t crop_size[3]) {
    clear_error();
    if ((!points_mm && npoints) || !frame_size || !frame_start_cm || !frame_spacing_cm || !crop_start || !crop_size)
        return set_error(CCCS_EINVAL, "null argument");
    // getROIExtents (DoseCalcIO/roi.cpp:245-316), including its FLT_MIN start of the maxima
    float mn[3], mx[3];
    for (int a = 0; a < 3; a++) {
        mn[a] = std::numeric_limits<float>::max();
        mx[a] = std::numeric_limits<float>::min();
    }
    for (uint64_t i = 0; i < npoints; i++)
        for (int a = 0; a < 3; a++) {
            const float v = points_mm[3 * i + a];
            if (v < mn[a]) mn[a] = v;
            if (v > mx[a]) mx[a] = v;
        }
    for (int a = 0; a < 3; a++) {
        int s = (int)std::floor((0.1f * mn[a] - frame_start_cm[a]) / frame_spacing_cm[a]);
        int e = (int)std::ceil((0.1f * mx[a] - frame_start_cm[a]) / frame_spacing_cm[a]);
        const int n = (int)frame_size[a];
        s = std::min(std::max(s, 0), n);
        e = std::min(std::max(e, 0), n);
        crop_start[a] = (uint32_t)s;
        crop_size[a] = (uint32_t)(e - s);
    }
    return CCCS_OK;
}

int cccs_roi_centroid(const uint8_t* mask, const uint32_t size[3], const float start_cm[3], const float spacing_cm[3],
                      float centroid_cm[3]) {
    clear_error();
    if (!mask || !size || !start_cm || !spacing_cm || !centroid_cm) return set_error(CCCS_EINVAL, "null argument");
    // getROICentroid (DoseCalcIO/roi.cpp:319-345): FP32 sums of voxel centres in voxel order
    float cx = 0.f, cy = 0.f, cz = 0.f, cnt = 0.f;
    const uint64_t nxy = (uint64_t)size[0] * size[1], n = nxy * size[2];
    for (uint64_t v = 0; v < n; v++) {
        if (!mask[v]) continue;
        const int X = (int)((v % nxy) % size[0]), Y = (int)((v % nxy) / size[0]), Z = (int)(v / nxy);
        cx += start_cm[0] + ((float)X + 0.5f) * spacing_cm[0];
        cy += start_cm[1] + ((float)Y + 0.5f) * spacing_cm[1];
        cz += start_cm[2] + ((float)Z + 0.5f) * spacing_cm[2];
        cnt += 1.f;
    }
    centroid_cm[0] = cx / cnt;
    centroid_cm[1] = cy / cnt;
    centroid_cm[2] = cz / cnt;
    return CCCS_OK;
}

int cccs_contour_rasterize(const float* points_mm, const uint32_t* sub_counts, int nsub, const uint32_t ct_size[3],
                           const float ct_start_cm[3], const float ct_spacing_cm[3], uint8_t* mask) {
    clear_error();
    if ((!points_mm && nsub > 0) || (!sub_counts && nsub > 0) || nsub < 0 || !ct_size || !ct_start_cm ||
        !ct_spacing_cm || !mask)
        return set_error(CCCS_EINVAL, "null argument");
    // per sub-contour: slice from the first point's z, vertices floor((x - start) / spacing), the
    // slice's polygon OR-ed into the mask (generateContourMask, cudaMemFunctions.cu:218-255)
    const int w = (int)ct_size[0], h = (int)ct_size[1];
    const uint64_t plane = (uint64_t)w * h;
    std::memset(mask, 0, plane * ct_size[2]);
    std::vector<uint8_t> img(plane);
    std::vector<int> xs, ys;
    uint64_t mark = 0;
    for (int c = 0; c < nsub; c++) {
        const uint32_t np = sub_counts[c];
        if (np == 0) continue;
        const float cz = 0.1f * points_mm[3 * mark + 2];
        const int z = (int)std::round((cz - ct_start_cm[2]) / ct_spacing_cm[2]);
        xs.resize(np);
        ys.resize(np);
        for (uint32_t p = 0; p < np; p++) {
            xs[p] = (int)std::floor((0.1f * points_mm[3 * (mark + p)] - ct_start_cm[0]) / ct_spacing_cm[0]);
            ys[p] = (int)std::floor((0.1f * points_mm[3 * (mark + p) + 1] - ct_start_cm[1]) / ct_spacing_cm[1]);
        }
        mark += np;
        if (z < 0 || z >= (int)ct_size[2]) continue;  // the reference writes outside its volume here
        std::fill(img.begin(), img.end(), 0);
        fill_polygon(img.data(), w, h, xs, ys);
        uint8_t* dst = mask + plane * z;
        for (uint64_t k = 0; k < plane; k++) dst[k] = (uint8_t)((img[k] + dst[k]) > 0);
    }
    return CCCS_OK;
}

int cccs_ct_lut_load(const char* path, float* hunits, float* massdens, int cap, int* n) {
    clear_error();
    if (!path || !n) return set_error(CCCS_EINVAL, "null argument");
    // load_lookup_table (DoseCalcIO/ctlut.cpp:39-90): "name CT# density [rel]" per line, '#'
    // comments and empty lines skipped; then sorted by CT number (CTLUT::sort)
    std::ifstream f(path);
    if (!f) return set_error(CCCS_EIO, std::string("cannot open ") + path);
    std::vector<std::pair<float, float>> pts;
    std::string line;
    int ln = 0;
    while (std::getline(f, line)) {
        ln++;
        size_t k = line.find_first_not_of(" \t\r");
        if (k == std::string::npos || line[k] == '#') continue;
        std::vector<std::string> fields;
        size_t a = 0;
        while (a < line.size()) {
            const size_t b = line.find(' ', a);
            const std::string t = line.substr(a, b == std::string::npos ? std::string::npos : b - a);
            if (!t.empty() && t != "\r") fields.push_back(t);
            if (b == std::string::npos) break;
            a = b + 1;
        }
        if (fields.size() < 3) return set_error(CCCS_EINVAL, "LUT line " + std::to_string(ln) + " has fewer than 3 fields");
        pts.emplace_back(std::strtof(fields[1].c_str(), nullptr), std::strtof(fields[2].c_str(), nullptr));
    }
    std::stable_sort(pts.begin(), pts.end(), [](const std::pair<float, float>& x, const std::pair<float, float>& y) {
        return x.first < y.first;
    });
    *n = (int)pts.size();
    if (!hunits || !massdens) return CCCS_OK;
    if (cap < (int)pts.size()) return set_error(CCCS_EINVAL, "LUT output too small");
    for (size_t i = 0; i < pts.size(); i++) {
        hunits[i] = pts[i].first;
        massdens[i] = pts[i].second;
    }
    return CCCS_OK;
}

}  // extern "C"
