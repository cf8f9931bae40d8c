// context.hip -- the resident device context (cccs_ctx), device buffers, per-kernel profiling,
// the line-table scan shared by the cone pipelines, and the context/memory C-ABI.
// Replaces the per-device static arrays, textures and memory accounting of initCudaConstandTex /
// freeCudaTexture (CCCS/dosecalc-beam/nvbbRayConvolve_host.cu:27-54, 158-293, 745-771;
// CCCS/CudaUtilities/memory_manager.cu:7-68).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"
#include "device_common.hpp"

namespace cccs {
// ------------------------------------------------------------------------------------------------
// buffers
int DevBuf::ensure(size_t need) {
    if (need <= bytes) return CCCS_OK;
    const size_t want = std::max(need, bytes + bytes / 4);
    if (p) {
        (void)hipDeviceSynchronize();  // queued kernels may still read the old buffer
        (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    if (hipMalloc(&p, want) != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        return set_error(CCCS_ENOMEM, "hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    bytes = want;
    return CCCS_OK;
}
void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}
int HostBuf::ensure(size_t need) {
    if (need <= bytes) return CCCS_OK;
    const size_t want = std::max(need, bytes + bytes / 4);
    if (p) {
        (void)hipDeviceSynchronize();  // queued copies may still use the old buffer
        (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        return set_error(CCCS_ENOMEM, "hipHostMalloc failed");
    }
    bytes = want;
    return CCCS_OK;
}
void HostBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

const char* const kKernelNames[K_COUNT] = {"terma", "crop", "cone", "pvcs", "bl_sample", "bl_siddon",
                                            "bl_walk", "bl_gather", "bl_unpack", "bl_compact", "plain_ray",
                                            "plain_conv", "terma_cached"};

static int prof_new_chunk(Profiler& P) {
    uint32_t* p = nullptr;
    if (hipMalloc((void**)&p, kCountChunk * sizeof(uint32_t)) != hipSuccess) {
        (void)hipGetLastError();
        return CCCS_ENOMEM;
    }
    P.chunks.push_back(p);
    return CCCS_OK;
}

uint32_t* prof_counter(cccs_ctx* c, size_t blocks) {
    Profiler& P = c->prof;
    P.pending = nullptr;
    P.pending_n = 0;
    if (!P.on || blocks == 0 || blocks > kCountChunk || P.chunks.empty()) return nullptr;
    if (P.chunk_used + blocks > kCountChunk) {
        P.cur++;
        P.chunk_used = 0;
        if (P.cur >= P.chunks.size() && prof_new_chunk(P) != CCCS_OK) {
            P.cur--;
            return nullptr;
        }
    }
    P.pending = P.chunks[P.cur] + P.chunk_used;
    P.pending_n = blocks;
    P.chunk_used += blocks;
    return P.pending;
}

int prof_begin(cccs_ctx* c, int kind) {
    Profiler& P = c->prof;
    if (!P.on) return -1;
    if (P.used + 2 > P.pool.size()) {
        for (int i = 0; i < 64; i++) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return -1;
            P.pool.push_back(e);
        }
    }
    ProfRec r;
    r.kind = kind;
    r.a = P.pool[P.used++];
    r.b = P.pool[P.used++];
    r.bytes = r.flops = 0;
    r.items = 0;
    r.cnt = P.pending;
    r.ncnt = P.pending_n;
    P.pending = nullptr;
    P.pending_n = 0;
    (void)hipEventRecord(r.a, c->stream);
    P.recs.push_back(r);
    return (int)P.recs.size() - 1;
}

void prof_end(cccs_ctx* c, int token, double bytes, double flops, uint64_t items) {
    if (token < 0) return;
    ProfRec& r = c->prof.recs[token];
    (void)hipEventRecord(r.b, c->stream);
    r.bytes = bytes;
    r.flops = flops;
    r.items = items;
}


thread_local std::string g_last_error;
int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
void clear_error() { g_last_error.clear(); }


// ---- line-table scan ---------------------------------------------------------------------------
constexpr int kScanBlock = 1024;

__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t* s) {
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        uint32_t add = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += add;
        __syncthreads();
    }
    return s[t];
}

__global__ void __launch_bounds__(kScanBlock) line_scan_local(LineMeta* __restrict__ meta, uint64_t n,
                                                            uint32_t* __restrict__ bsum) {
    __shared__ uint32_t s[kScanBlock];
    const uint64_t i = (uint64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const uint32_t len = i < n ? (uint32_t)meta[i].len : 0u;
    const uint32_t inc = block_inclusive_scan(len, s);
    if (i < n) meta[i].off = inc - len;
    if (threadIdx.x == kScanBlock - 1) bsum[blockIdx.x] = inc;
}

__global__ void __launch_bounds__(kScanBlock) line_scan_blocks(uint32_t* __restrict__ bsum, uint32_t nb) {
    __shared__ uint32_t s[kScanBlock];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += kScanBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? bsum[i] : 0u;
        const uint32_t inc = block_inclusive_scan(v, s);
        if (i < nb) bsum[i] = carry + inc - v;
        carry += s[kScanBlock - 1];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kScanBlock) line_scan_add(LineMeta* __restrict__ meta, uint64_t n,
                                                          const uint32_t* __restrict__ bsum) {
    const uint64_t i = (uint64_t)blockIdx.x * kScanBlock + threadIdx.x;
    if (i < n) meta[i].off += bsum[blockIdx.x];
}

size_t line_scan_words(uint64_t n) { return (size_t)((n + kScanBlock - 1) / kScanBlock); }

int launch_line_scan(cccs_ctx* c, LineMeta* meta, uint64_t n) {
    if (n == 0) return CCCS_OK;
    CCCS_TRY(c->scan.ensure(sizeof(uint32_t) * line_scan_words(n)));
    return launch_line_scan(c, meta, n, c->scan.as<uint32_t>());
}

int launch_line_scan(cccs_ctx* c, LineMeta* meta, uint64_t n, uint32_t* bsum) {
    if (n == 0) return CCCS_OK;
    const uint64_t nb = line_scan_words(n);
    if (nb > 0xffffffffULL) return set_error(CCCS_EGEOM, "too many REV lines");
    hipLaunchKernelGGL(line_scan_local, dim3((unsigned)nb), dim3(kScanBlock), 0, c->stream, meta, n, bsum);
    hipLaunchKernelGGL(line_scan_blocks, dim3(1), dim3(kScanBlock), 0, c->stream, bsum, (uint32_t)nb);
    hipLaunchKernelGGL(line_scan_add, dim3((unsigned)nb), dim3(kScanBlock), 0, c->stream, meta, n, bsum);
    CCCS_HIP(hipGetLastError());
    return CCCS_OK;
}


}  // namespace cccs

using namespace cccs;

// =================================================================================================
// C-ABI

extern "C" const char* cccs_last_error(void) { return g_last_error.c_str(); }

extern "C" int cccs_ctx_create(const cccs_problem* p, int device, cccs_ctx** out) {
    clear_error();
    if (!p || !out) return set_error(CCCS_EINVAL, "null argument");
    *out = nullptr;
    if (!p->density || !p->kernel || !p->radii || !p->conv_theta_deg || !p->conv_phi_deg || !p->spec_fluence ||
        !p->spec_energy || !p->spec_mu || !p->spec_mu_en)
        return set_error(CCCS_EINVAL, "problem tables must be non-null");
    if (p->ntheta < 2 || p->ntheta % 2 || p->nphi < 1 || p->nradii < 2 || p->nradii > (uint32_t)kMaxRadii)
        return set_error(CCCS_EINVAL, "bad ntheta/nphi/nradii");
    if (p->nphi * (p->ntheta / 2) > (uint32_t)kMaxRays) return set_error(CCCS_EINVAL, "more than 48 cone lines");
    if (p->nspec < 1 || p->nspec > (uint32_t)kMaxSpec) return set_error(CCCS_EINVAL, "nspec must be in [1, 30]");
    for (int k = 0; k < 3; k++) {
        if (p->size[k] == 0 || p->bbox_size[k] == 0 || (uint64_t)p->bbox_start[k] + p->bbox_size[k] > p->size[k])
            return set_error(CCCS_EINVAL, "calc_bbox must lie inside the volume");
        if (!(p->voxel[k] > 0.f)) return set_error(CCCS_EINVAL, "voxel size must be > 0");
    }
    if (!(p->rev_latspacing > 0.f) || !(p->rev_longspacing > 0.f)) return set_error(CCCS_EINVAL, "REV spacing must be > 0");
    for (uint32_t r = 1; r < p->nradii; r++)
        if (!(p->radii[r] > p->radii[r - 1]) || !(p->radii[0] > 0.f)) return set_error(CCCS_EINVAL, "radii must increase");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return set_error(CCCS_EDEVICE, "no HIP device available");
    }
    if (device < 0 || device >= ndev) return set_error(CCCS_EINVAL, "device ordinal out of range");
    CCCS_HIP(hipSetDevice(device));
    cccs_ctx* c = new cccs_ctx();
    c->device = device;
    c->prob = *p;
    c->prob.density = nullptr; c->prob.kernel = nullptr; c->prob.radii = nullptr;
    c->prob.conv_theta_deg = nullptr; c->prob.conv_phi_deg = nullptr;
    c->prob.spec_fluence = c->prob.spec_energy = c->prob.spec_mu = c->prob.spec_mu_en = nullptr;
    c->nrays = (int)(p->nphi * (p->ntheta / 2));
    c->kernel.assign(p->kernel, p->kernel + (size_t)p->ntheta * p->nradii);
    c->radii.assign(p->radii, p->radii + p->nradii);
    c->theta.assign(p->conv_theta_deg, p->conv_theta_deg + c->nrays);
    c->phi.assign(p->conv_phi_deg, p->conv_phi_deg + c->nrays);
    c->vol.nx = (int)p->size[0]; c->vol.ny = (int)p->size[1]; c->vol.nz = (int)p->size[2];
    c->vol.start = ld3(p->start);
    c->vol.voxel = ld3(p->voxel);
    c->vol.end = mk3(p->start[0] + (float)p->size[0] * p->voxel[0], p->start[1] + (float)p->size[1] * p->voxel[1],
                     p->start[2] + (float)p->size[2] * p->voxel[2]);
    c->spec.n = (int)p->nspec;
    for (uint32_t e = 0; e < p->nspec; e++) {
        c->spec.flu[e] = p->spec_fluence[e];
        c->spec.en[e] = p->spec_energy[e];
        c->spec.mu[e] = p->spec_mu[e];
    }
    // beam-path REV limit, derived as initCudaConstandTex does (nvbbRayConvolve_host.cu:248-258)
    {
        float min_vox = std::min(p->rev_latspacing, p->rev_longspacing);
        v3 bb = ldu3(p->bbox_size);
        unsigned diag = (unsigned)std::ceil(len3(bb) * p->voxel[0] / min_vox);
        unsigned v = (unsigned)(diag + 2 * (0.0f / min_vox + p->penumbra / min_vox) + 100);
        c->max_rev_beam[0] = c->max_rev_beam[1] = c->max_rev_beam[2] = v;
    }
    int st = CCCS_OK;
    do {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { st = set_error(CCCS_EDEVICE, "stream"); break; }
        bool sok = true;
        for (hipStream_t& x : c->xstream) sok = sok && hipStreamCreateWithFlags(&x, hipStreamNonBlocking) == hipSuccess;
        if (!sok) { st = set_error(CCCS_EDEVICE, "stream"); break; }
        bool evok = true;
        for (hipEvent_t& e : c->ray_ev) evok = evok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
        if (!evok) { st = set_error(CCCS_EDEVICE, "events"); break; }
        size_t nv = (size_t)p->size[0] * p->size[1] * p->size[2];
        if (hipMalloc((void**)&c->d_dens, nv * sizeof(float)) != hipSuccess) { st = set_error(CCCS_ENOMEM, "density"); break; }
        if (hipMemcpy(c->d_dens, p->density, nv * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) { st = set_error(CCCS_EDEVICE, "density copy"); break; }
        // tables: kernel | radii | 1/width | mean radii | 1/width of mean radii
        const uint32_t nr = p->nradii;
        std::vector<float> tab((size_t)p->ntheta * nr + 4 * nr);
        std::copy(c->kernel.begin(), c->kernel.end(), tab.begin());
        float* R = tab.data() + (size_t)p->ntheta * nr;
        float* IW = R + nr;
        float* MR = IW + nr;
        float* MIW = MR + nr;
        for (uint32_t r = 0; r < nr; r++) {
            R[r] = p->radii[r];
            IW[r] = 1.0f / (r == 0 ? p->radii[0] : (p->radii[r] - p->radii[r - 1]));
            MR[r] = (float)(r == 0 ? 0.5 * p->radii[0] : 0.5 * (p->radii[r] + p->radii[r - 1]));
        }
        for (uint32_t r = 0; r < nr; r++) MIW[r] = 1.0f / (r == 0 ? MR[0] : (MR[r] - MR[r - 1]));
        if (hipMalloc((void**)&c->d_tables, tab.size() * sizeof(float)) != hipSuccess) { st = set_error(CCCS_ENOMEM, "tables"); break; }
        if (hipMemcpy(c->d_tables, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) { st = set_error(CCCS_EDEVICE, "tables copy"); break; }
        c->resident = nv * sizeof(float) + tab.size() * sizeof(float);
        uint64_t h = 1469598103934665603ull;
        const uint32_t* bits = reinterpret_cast<const uint32_t*>(p->density);
        for (size_t i = 0; i < nv; i++) h = (h ^ bits[i]) * 1099511628211ull;
        c->dens_hash = h;
        // the no-allocation entry point's fixed scratch: its address on this device, resolved once
        if ((st = noalloc_scratch_init(device)) != CCCS_OK) break;
    } while (false);
    if (st != CCCS_OK) {
        cccs_ctx_destroy(c);
        return st;
    }
    *out = c;
    return CCCS_OK;
}

extern "C" int cccs_ctx_destroy(cccs_ctx* c) {
    if (!c) return CCCS_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipStream_t x : c->xstream)
        if (x) (void)hipStreamSynchronize(x);
    if (c->cstream) (void)hipStreamSynchronize(c->cstream);
    if (c->d_dens) (void)hipFree(c->d_dens);
    if (c->d_tables) (void)hipFree(c->d_tables);
    c->terma.release(); c->fmap.release(); c->partial.release(); c->rev.release(); c->cones.release();
    c->dose.release(); c->col.release(); c->lines.release(); c->scan.release(); c->dt.release();
    c->cls_cnt.release(); c->cls_list.release();
    for (int i = 0; i < kRaySets - 1; i++) {
        c->xlines[i].release(); c->xrev[i].release(); c->xh[i].release(); c->xqueue[i].release(); c->xscan[i].release();
        c->xlcone[i].release();
    }
    c->h_small.release(); c->h_cones.release(); c->h_stage.release();
    for (hipEvent_t& e : c->stage_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->copy_ev) (void)hipEventDestroy(c->copy_ev);
    if (c->cstream) (void)hipStreamDestroy(c->cstream);
    c->bl_tab.release(); c->bl_pdose.release(); c->bl_lcone.release(); c->bl_rowcnt.release(); c->bl_arena.release();
    c->bl_cnt.release(); c->bl_keys.release(); c->bl_vals.release(); c->bl_fmap.release();
    c->bl_h.release(); c->bl_queue.release(); c->bl_qcnt.release(); c->bl_tiles.release();
    c->h_bl.release(); c->h_bl_out.release(); c->h_bl_out2.release();
    c->roi_tab.release(); c->roi_rank.release(); c->roi_cnt.release(); c->roi_keys.release(); c->roi_vals.release();
    for (hipEvent_t e : c->prof.pool) (void)hipEventDestroy(e);
    for (uint32_t* p : c->prof.chunks) (void)hipFree(p);
    for (hipEvent_t e : c->ray_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t x : c->xstream)
        if (x) (void)hipStreamDestroy(x);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CCCS_OK;
}

extern "C" int cccs_ctx_memory(const cccs_ctx* c, uint64_t* resident, uint64_t* scratch) {
    if (!c) return CCCS_EINVAL;
    if (resident) *resident = c->resident;
    if (scratch)
        *scratch = c->terma.bytes + c->fmap.bytes + c->partial.bytes + c->rev.bytes + c->cones.bytes + c->dose.bytes +
                   c->col.bytes + c->dt.bytes;
    return CCCS_OK;
}

extern "C" int cccs_ctx_set_schedule(cccs_ctx* c, const char* key, int value) {
    clear_error();
    if (!c || !key) return set_error(CCCS_EINVAL, "null argument");
    if (std::strcmp(key, "pvcs_pass_cones") == 0) {
        if (value < 0 || value > kMaxRays) return set_error(CCCS_EINVAL, "pvcs_pass_cones must be in [0, 48]");
        c->sched.pvcs_pass_cones = value;
        return CCCS_OK;
    }
    if (std::strcmp(key, "cone_ramp_min") == 0) {
        if (value < 1) return set_error(CCCS_EINVAL, "cone_ramp_min must be >= 1");
        c->sched.cone_ramp_min = value;
        return CCCS_OK;
    }
    if (std::strcmp(key, "bl_job_gb") == 0) {
        if (value < 1 || value > 1024) return set_error(CCCS_EINVAL, "bl_job_gb must be in [1, 1024]");
        c->sched.bl_job_gb = value;
        return CCCS_OK;
    }
    if (std::strcmp(key, "bl_streams") == 0) {
        if (value < 1 || value > kRaySets) return set_error(CCCS_EINVAL, "bl_streams must be in [1, 4]");
        c->sched.bl_streams = value;
        return CCCS_OK;
    }
    if (std::strcmp(key, "bl_walk_split") == 0) {
        if (value < 0 || value > kMaxLine) return set_error(CCCS_EINVAL, "bl_walk_split must be in [0, 1024]");
        c->sched.bl_walk_split = value;
        return CCCS_OK;
    }
    if (std::strcmp(key, "bl_sample_waves") == 0 || std::strcmp(key, "bl_walk_waves") == 0) {
        if (value < 64 || value > (1 << 22)) return set_error(CCCS_EINVAL, "wave count must be in [64, 2^22]");
        (key[3] == 's' ? c->sched.bl_sample_waves : c->sched.bl_walk_waves) = value;
        return CCCS_OK;
    }
    for (int i = 0; i < kConeClasses; i++) {
        const std::string k = "cone_class_threads" + std::to_string(i);
        if (k == key) {
            if (value != 64 && value != 128 && value != 256) return set_error(CCCS_EINVAL, "threads must be 64, 128 or 256");
            c->sched.cone_class_threads[i] = value;
            return CCCS_OK;
        }
    }
    for (int i = 0; i < kConeClasses - 1; i++) {
        const std::string k = "cone_class_len" + std::to_string(i);
        if (k == key) {
            if (value < 1 || value > kMaxLine) return set_error(CCCS_EINVAL, "class length out of range");
            c->sched.cone_class_len[i] = value;
            return CCCS_OK;
        }
    }
    return set_error(CCCS_EINVAL, std::string("unknown schedule key: ") + key);
}

extern "C" int cccs_ctx_get_schedule(const cccs_ctx* c, const char* key, int* value) {
    clear_error();
    if (!c || !key || !value) return set_error(CCCS_EINVAL, "null argument");
    const Schedule& S = c->sched;
    if (std::strcmp(key, "cone_ramp_min") == 0) { *value = S.cone_ramp_min; return CCCS_OK; }
    if (std::strcmp(key, "pvcs_pass_cones") == 0) { *value = S.pvcs_pass_cones; return CCCS_OK; }
    if (std::strcmp(key, "bl_sample_waves") == 0) { *value = S.bl_sample_waves; return CCCS_OK; }
    if (std::strcmp(key, "bl_walk_waves") == 0) { *value = S.bl_walk_waves; return CCCS_OK; }
    if (std::strcmp(key, "bl_streams") == 0) { *value = S.bl_streams; return CCCS_OK; }
    if (std::strcmp(key, "bl_job_gb") == 0) { *value = S.bl_job_gb; return CCCS_OK; }
    if (std::strcmp(key, "bl_walk_split") == 0) { *value = S.bl_walk_split; return CCCS_OK; }
    for (int i = 0; i < kConeClasses; i++) {
        if (("cone_class_threads" + std::to_string(i)) == key) { *value = S.cone_class_threads[i]; return CCCS_OK; }
        if (i < kConeClasses - 1 && ("cone_class_len" + std::to_string(i)) == key) {
            *value = S.cone_class_len[i];
            return CCCS_OK;
        }
    }
    return set_error(CCCS_EINVAL, std::string("unknown schedule key: ") + key);
}

// ---- memory helpers ----------------------------------------------------------------------------
extern "C" int cccs_dev_alloc(cccs_ctx* ctx, size_t bytes, void** ptr) {
    clear_error();
    if (!ctx || !ptr) return set_error(CCCS_EINVAL, "null argument");
    *ptr = nullptr;
    CCCS_HIP(hipSetDevice(ctx->device));
    if (hipMalloc(ptr, std::max<size_t>(bytes, 4)) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return set_error(CCCS_ENOMEM, "hipMalloc failed");
    }
    return CCCS_OK;
}
extern "C" int cccs_dev_free(cccs_ctx* ctx, void* ptr) {
    if (!ctx) return CCCS_EINVAL;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ptr) CCCS_HIP(hipFree(ptr));
    return CCCS_OK;
}
extern "C" int cccs_dev_upload(cccs_ctx* ctx, void* dst, const void* src, size_t bytes) {
    clear_error();
    if (!ctx || !dst || !src) return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    CCCS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return CCCS_OK;
}
namespace cccs {
int download_staged(cccs_ctx* c, void* dst, const void* src, size_t bytes, bool add_f32) {
    constexpr size_t kChunk = size_t(8) << 20;
    CCCS_TRY(c->h_stage.ensure(2 * kChunk));
    for (hipEvent_t& e : c->stage_ev)
        if (!e) CCCS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    char* stage = c->h_stage.as<char>();
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    const size_t n = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](size_t i) -> int {
        const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
        CCCS_HIP(hipMemcpyAsync(stage + (i & 1) * kChunk, s + off, len, hipMemcpyDeviceToHost, c->stream));
        CCCS_HIP(hipEventRecord(c->stage_ev[i & 1], c->stream));
        return CCCS_OK;
    };
    if (n > 0) CCCS_TRY(issue(0));
    for (size_t i = 0; i < n; i++) {
        if (i + 1 < n) CCCS_TRY(issue(i + 1));
        CCCS_HIP(hipEventSynchronize(c->stage_ev[i & 1]));
        const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
        const char* from = stage + (i & 1) * kChunk;
        if (add_f32) {
            float* o = reinterpret_cast<float*>(d + off);
            const float* f = reinterpret_cast<const float*>(from);
            for (size_t k = 0; k < len / sizeof(float); k++) o[k] += f[k];
        } else {
            std::memcpy(d + off, from, len);
        }
    }
    return CCCS_OK;
}
}  // namespace cccs

extern "C" int cccs_dev_download(cccs_ctx* ctx, void* dst, const void* src, size_t bytes) {
    clear_error();
    if (!ctx || !dst || !src) return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    return download_staged(ctx, dst, src, bytes, false);
}
extern "C" int cccs_dev_zero(cccs_ctx* ctx, void* ptr, size_t bytes) {
    clear_error();
    if (!ctx || !ptr) return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    CCCS_HIP(hipMemsetAsync(ptr, 0, bytes, ctx->stream));
    return CCCS_OK;
}
__global__ void __launch_bounds__(256) copy_kernel(float4* __restrict__ dst, const float4* __restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

extern "C" int cccs_dev_copy(cccs_ctx* ctx, void* dst, const void* src, size_t bytes) {
    clear_error();
    if (!ctx || !dst || !src || bytes % 16 != 0) return set_error(CCCS_EINVAL, "bad copy arguments");
    CCCS_HIP(hipSetDevice(ctx->device));
    size_t n = bytes / 16;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    if (n) hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, ctx->stream, (float4*)dst, (const float4*)src, n);
    CCCS_HIP(hipGetLastError());
    return CCCS_OK;
}

extern "C" int cccs_ctx_sync(cccs_ctx* ctx) {
    clear_error();
    if (!ctx) return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    for (hipStream_t x : ctx->xstream) CCCS_HIP(hipStreamSynchronize(x));
    if (ctx->cstream) CCCS_HIP(hipStreamSynchronize(ctx->cstream));
    return CCCS_OK;
}

// ---- profiling -----------------------------------------------------------------------------------
// Algorithmic FP32 operations per counted unit: one line-walk iteration (dens*step, two half-step
// adds, interval lerp, kernel difference times TERMA: 10 ops) and one Siddon plane crossing (three
// validity tests, plane choice, segment length, midpoint fetch coordinate, 8-tap trilinear: 40 ops).
static const double kFlopsPerWalkIter = 10.0;
static const double kFlopsPerSiddonStep = 40.0;
static const double kBytesPerStoredSample = 16.0;

extern "C" int cccs_ctx_profile(cccs_ctx* ctx, int enable) {
    clear_error();
    if (!ctx) return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    Profiler& P = ctx->prof;
    // allocate the first counter chunks now, outside any timed region
    while (enable && P.chunks.size() < 2)
        if (prof_new_chunk(P) != CCCS_OK) return set_error(CCCS_ENOMEM, "profiling counters");
    P.on = enable != 0;
    return CCCS_OK;
}

extern "C" int cccs_ctx_profile_reset(cccs_ctx* ctx) {
    clear_error();
    if (!ctx) return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    Profiler& P = ctx->prof;
    P.recs.clear();
    P.used = 0;
    P.cur = 0;  // counter chunks stay allocated and are refilled from the first one
    P.chunk_used = 0;
    return CCCS_OK;
}

extern "C" int cccs_ctx_profile_read(cccs_ctx* ctx, cccs_kernel_stats* out, int max_rows, int* nrows) {
    clear_error();
    if (!ctx || !out || !nrows) return set_error(CCCS_EINVAL, "null argument");
    CCCS_HIP(hipSetDevice(ctx->device));
    CCCS_HIP(hipStreamSynchronize(ctx->stream));
    cccs_kernel_stats rows[K_COUNT];
    std::memset(rows, 0, sizeof(rows));
    uint64_t cnt[K_COUNT] = {0};
    std::vector<uint32_t> tmp;
    for (const ProfRec& r : ctx->prof.recs) {
        float ms = 0.f;
        CCCS_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        cccs_kernel_stats& k = rows[r.kind];
        k.launches++;
        k.ms_total += ms;
        k.alg_bytes += r.bytes;
        k.alg_flops += r.flops;
        k.work_items += r.items;
        if (r.cnt && r.ncnt) {
            tmp.resize(r.ncnt);
            CCCS_HIP(hipMemcpy(tmp.data(), r.cnt, r.ncnt * sizeof(uint32_t), hipMemcpyDeviceToHost));
            for (uint32_t v : tmp) cnt[r.kind] += v;
        }
    }
    rows[K_CONE].alg_flops += kFlopsPerWalkIter * (double)cnt[K_CONE];
    rows[K_BL_WALK].alg_flops += kFlopsPerWalkIter * (double)cnt[K_BL_WALK];
    rows[K_TERMA].alg_flops += kFlopsPerSiddonStep * (double)cnt[K_TERMA];
    rows[K_BL_SIDDON].alg_flops += kFlopsPerSiddonStep * (double)cnt[K_BL_SIDDON];
    // the sample stage counts stored REV samples and is priced as the BEV resampling of a cone:
    // density and TERMA read at each sample (8 B) and written line-major (8 B)
    rows[K_BL_SAMPLE].alg_bytes += kBytesPerStoredSample * (double)cnt[K_BL_SAMPLE];
    rows[K_BL_SAMPLE].work_items = cnt[K_BL_SAMPLE];
    int n = 0;
    for (int k = 0; k < K_COUNT && n < max_rows; k++) {
        if (rows[k].launches == 0) continue;
        std::snprintf(rows[k].name, sizeof(rows[k].name), "%s", kKernelNames[k]);
        out[n++] = rows[k];
    }
    *nrows = n;
    return CCCS_OK;
}
