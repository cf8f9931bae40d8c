// context.hpp -- the resident device context behind cccs_ctx (replaces the per-device static
// arrays and textures of initCudaConstandTex, dosecalc-beam/nvbbRayConvolve_host.cu:27-54).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "cccs_internal.hpp"
#include "errors.hpp"

namespace cccs {

// Grow-only device buffer; contents are not preserved across growth.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Grow-only pinned host buffer.
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

constexpr int kMaxRays = 48;      // ntheta/2 * nphi <= 48 cone lines per beam in the plain path
constexpr int kMaxLine = 1024;    // REV line limit of the reference (nvbbRayConvolve_host.cu:567)
constexpr int kMaxRadii = 32;

// Per-cone constants of the plain beam path, consumed by the cone and PVCS kernels.
struct ConeDev {
    m33 kinv;          // BEV <- REV (index units)
    m33 a;             // REV <- PVCS index: kernel_fwd * beam_fwd
    v3 rs;             // REV origin in REV index units (revStart / rev spacing)
    v3 t;              // REV coordinate of PVCS index (0,0,0)
    int nx, ny, nz;    // line length, lines along y, lines along z
    int rowf, rowb;    // kernel rows of the forward / backward walk
    uint64_t line0;    // first global line index of this cone
    uint64_t off;      // float offset of this cone's REV dose grid
};

struct BeamGeom {
    v3 lim_min;            // BEV crop-box origin in PVCS [cm]
    uint32_t bev_size[3];
    v3 bev_min, bev_max;
    int nrays;
    v3 rev_min[kMaxRays];
    uint32_t rev_size[kMaxRays][3];
    uint64_t total_lines, total_rev;
    int pvcs_lo[3], pvcs_hi[3];  // PVCS index box that can receive dose (half-open)
};

// ---- profiling ---------------------------------------------------------------------------------
enum KernelKind { K_TERMA = 0, K_CROP, K_CONE, K_PVCS, K_BL_SAMPLE, K_BL_SIDDON, K_BL_WALK, K_BL_GATHER,
                  K_BL_UNPACK, K_BL_COMPACT, K_PL_RAY, K_PL_CONV, K_TERMA_CACHED, K_COUNT };
extern const char* const kKernelNames[K_COUNT];

struct ProfRec {
    int kind;
    hipEvent_t a, b;
    double bytes, flops;
    uint64_t items;
    uint32_t* cnt;  // per-workgroup work counts of this launch (device), or nullptr
    size_t ncnt;
};

constexpr size_t kCountChunk = size_t(1) << 24;  // counter entries per device chunk (64 MiB)

struct Profiler {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<ProfRec> recs;
    std::vector<uint32_t*> chunks;  // per-launch counter slices, summed on the host at read time
    size_t cur = 0;                 // chunk being filled
    size_t chunk_used = 0;
    uint32_t* pending = nullptr;
    size_t pending_n = 0;
};

struct LineMeta;
}  // namespace cccs

namespace cccs {
constexpr int kConeClasses = 4;
constexpr int kRaySets = 4;  // most cone directions of the beamlet pipeline in flight
// Launch-schedule knobs (cccs_ctx_set_schedule).  They change how the work is scheduled, never
// which samples contribute.
struct Schedule {
    // plain-path cone lines whose valid segment has at least this many samples use the
    // segment-sum walk (O(nradii) per destination) instead of the sequential walk (O(line))
    int cone_ramp_min = 64;
    // cones per pass of the PVCS resample (each pass adds its cones' dose to the volume); 0: from
    // the size of the beam's REV data
    int pvcs_pass_cones = 0;
    // upper stored-line lengths of the cone-kernel launch classes (the last class takes the rest)
    int cone_class_len[kConeClasses] = {96, 160, 320, kMaxLine};
    // threads per line (64, 128 or 256) of each class: long lines spread over more waves
    int cone_class_threads[kConeClasses] = {64, 64, 128, 128};
    // waves of the beamlet path's persistent per-line kernels (each takes a stride of lines).  The
    // sample stage holds 3 waves per SIMD: with several cone directions in flight, the other
    // streams' Siddon, walk and gather waves run beside it (its own 5-wave occupancy would take
    // every register of the SIMD)
    int bl_sample_waves = 3072;
    int bl_walk_waves = 32768;
    // device memory of one beamlet job [GiB]: batches join a job while its packed grids, output
    // bounds and per-direction buffers fit; fewer, larger jobs drain the direction pipeline less often
    int bl_job_gb = 12;
    // the beamlet walk sizes its LDS for the longest line of its launch; it runs one launch per
    // length class: lines up to this many stored samples, up to twice that, ..., and the rest
    // (0: one launch).  With four directions in flight, 128 beat 96, 160 and 256 (1.2 % on the step)
    int bl_walk_split = 128;
    // cone directions of the beamlet path in flight at once (1 .. kRaySets): direction r runs on
    // stream r mod bl_streams, so consecutive directions' sample, Siddon and walk overlap; the
    // gathers stay in direction order (measured on the 50-beam job: 1 -> 10.8k, 2 -> 12.3k,
    // 3 -> 12.6k, 4 -> 13.0k beamlet-doses/s)
    int bl_streams = 4;
};
}  // namespace cccs

struct cccs_ctx {
    int device = 0;
    uint64_t dens_hash = 0;  // FNV-1a of the density values (matches an angle cache to its problem)
    const struct cccs_angle_cache* angle_cache = nullptr;  // resident unit TERMA per beam source
    hipStream_t stream = nullptr;
    // extra streams of the beamlet path's multi-direction pipeline, and its events (one per set
    // plus a start event; created with the context, default priority like the first)
    hipStream_t xstream[cccs::kRaySets - 1] = {};
    hipEvent_t ray_ev[cccs::kRaySets + 1] = {};
    cccs_problem prob{};  // scalar copy (table pointers cleared)
    cccs::VolDesc vol{};
    cccs::SpecDev spec{};
    std::vector<float> kernel, radii, theta, phi;
    uint32_t max_rev_beam[3] = {0, 0, 0};
    int nrays = 0;
    float* d_dens = nullptr;
    float* d_tables = nullptr;  // kernel (ntheta*nradii) | radii | inverse widths | mean radii | inverse widths
    size_t resident = 0;
    cccs::DevBuf terma, fmap, partial, rev, cones, dose, col, lines, scan, dt;
    cccs::DevBuf cls_cnt, cls_list;  // plain path: per-block class counts/offsets, per-class line lists
    // ray sets 1.. of the beamlet pipeline (set 0 is lines/rev/bl_h/bl_queue/scan/bl_lcone): line
    // table, REV store, half steps, TERMA queue, scan partials, grid of every line
    cccs::DevBuf xlines[cccs::kRaySets - 1], xrev[cccs::kRaySets - 1], xh[cccs::kRaySets - 1],
        xqueue[cccs::kRaySets - 1], xscan[cccs::kRaySets - 1], xlcone[cccs::kRaySets - 1];
    cccs::HostBuf h_small, h_cones, h_stage;  // h_stage: pinned double buffer of download_staged
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    hipStream_t cstream = nullptr;  // beamlet columns to pinned host memory, beside the next job
    hipEvent_t copy_ev = nullptr;
    // beamlet path: job tables, packed grids, column arena, sparse output, line grids
    cccs::DevBuf bl_tab, bl_pdose, bl_arena, bl_cnt, bl_keys, bl_vals, bl_fmap, bl_h, bl_queue, bl_qcnt, bl_tiles,
        bl_lcone, bl_rowcnt;
    cccs::HostBuf h_bl, h_bl_out, h_bl_out2;  // h_bl_out/h_bl_out2: sparse columns of alternate jobs
    std::vector<uint8_t> scratch_host;
    cccs::Profiler prof;
    cccs::Schedule sched;
    // ROI reduction (cccs_set_roi_masks): per-ROI table, rank volumes over the crop boxes, and the
    // reduced output of the current job
    int nroi = 0;
    cccs::DevBuf roi_tab, roi_rank, roi_cnt, roi_keys, roi_vals;
};

namespace cccs {
// table offsets inside d_tables
inline size_t tab_kernel(const cccs_ctx*) { return 0; }
inline size_t tab_radii(const cccs_ctx* c) { return (size_t)c->prob.ntheta * c->prob.nradii; }
inline size_t tab_iw(const cccs_ctx* c) { return tab_radii(c) + c->prob.nradii; }
inline size_t tab_mradii(const cccs_ctx* c) { return tab_iw(c) + c->prob.nradii; }
inline size_t tab_miw(const cccs_ctx* c) { return tab_mradii(c) + c->prob.nradii; }

BeamDev make_beam_dev(const cccs_beam& b);
// Profiling hooks.  prof_counter reserves a per-workgroup count slice for the next launch (nullptr
// when profiling is off); prof_begin records a start event and attaches that slice; prof_end
// records the stop event and the launch's algorithmic bytes/FLOPs.  Nothing runs in the stream
// besides the events: the counts are summed on the host by cccs_ctx_profile_read.
uint32_t* prof_counter(cccs_ctx* c, size_t blocks);
int prof_begin(cccs_ctx* c, int kind);
void prof_end(cccs_ctx* c, int token, double bytes, double flops, uint64_t items);

// Device-to-host copy of `bytes` (a multiple of 4) through a pinned double buffer: the DMA of one
// chunk overlaps the host copy (or FP32 add, when add_f32) of the previous one.  Waits for the
// context stream first.
int download_staged(cccs_ctx* c, void* dst, const void* src, size_t bytes, bool add_f32);

// Exclusive scan of LineMeta::len into LineMeta::off over n lines (on the context stream), with
// the block partials in the context's scan buffer, or in `bsum` (line_scan_words(n) words).
int launch_line_scan(cccs_ctx* c, LineMeta* meta, uint64_t n);
int launch_line_scan(cccs_ctx* c, LineMeta* meta, uint64_t n, uint32_t* bsum);
size_t line_scan_words(uint64_t n);
int check_beam(const cccs_ctx* ctx, const cccs_beam& b);
// Resolves, once per device and with that device current, the address of the fixed scratch array
// used by cccs_beam_dose_device_noalloc.
int noalloc_scratch_init(int device);
}  // namespace cccs
