// server.cpp -- request loop for single-beam dose and beamlet-column requests, the calls a
// beam-orientation optimiser issues while it anneals.
//
// * Resident state: each slot owns a context (density, kernel tables and cone angles stay on
//   the device) with its own stream.  Slots run concurrently, so their work overlaps on the GPU.
// * Continuous batching: a free slot takes every queued beamlet request that fits its share, not
//   just one, and runs them as one multi-beam job (shared launches).  Requests that arrive while
//   a batch runs join the next batch.  A full-beam request fills the device alone and shares no
//   launches, so a slot takes one at a time; concurrent slots overlap them.
// * Latency first: queued full-beam requests are taken before queued beamlet requests, whatever
//   their arrival order, and with two or more slots at most all but one slot run beamlet batches
//   at a time, so a full-beam request never waits for a long beamlet batch (a precompute that
//   shares the resident state with an optimiser's requests).  One beamlet batch already keeps
//   the device busy with its several cone directions in flight.
// * Admission by HBM footprint: a request is admitted only while the device bytes in flight plus
//   its estimate stay within the budget.  A request larger than the whole budget runs alone.
// * Latency: each request's submit-to-completion time.  Statistics give p50/p99.
// Host code only; it drives the library through its C entry points.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cccs.h"
#include "errors.hpp"

using namespace cccs;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    int kind = 0;
    cccs_beam beam{};
    std::vector<float> fluence;
    float threshold = 0.f;
    int mode = 0;
    uint64_t bytes = 0;  // admission estimate
    Clock::time_point submitted, finished;
    bool done = false;
    int status = CCCS_OK;
    std::string error;
    int batch_size = 0;
    std::vector<float> dose;  // beam-dose result, full volume
    // beamlet result: columns back to back
    std::vector<uint64_t> keys;
    std::vector<float> values;
    std::vector<cccs_sparse_column> cols;  // keys/values fields hold offsets until delivery
};

}  // namespace

struct cccs_server {
    std::vector<cccs_ctx*> slots;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> q_dose, q_bl;  // queued full-beam and beamlet requests
    int bl_running = 0;                                // slots running a beamlet batch
    int running = 0;                                   // slots running any batch
    cccs_angle_cache* angles = nullptr;                // resident unit TERMA shared by every slot
    std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
    uint64_t next_id = 1;
    uint64_t budget = 0, in_flight = 0;
    int max_batch = 64;
    bool stop = false;
    size_t nvox = 0, nbbox = 0;
    std::vector<double> latencies;
    uint64_t batches = 0, batched = 0;
    // full-volume result buffers returned by cccs_server_wait, reused so that a request does not
    // pay for first-touch page faults of a fresh 4 * nvox byte allocation
    std::vector<std::vector<float>> dose_pool;
};

namespace {

// Device bytes a request may need while it runs: for a beam dose the TERMA over the calc_bbox,
// the compact REV store (bounded by a few volumes) and the dose volume; for a beamlet request the
// library's own per-batch footprint.
uint64_t estimate_bytes(cccs_server* s, cccs_ctx* ctx, const Request& r) {
    if (r.kind == CCCS_REQ_BEAMLET) {
        uint64_t b = 0;
        if (cccs_beamlet_footprint(ctx, &r.beam, r.fluence.data(), &b) == CCCS_OK) return b;
        clear_error();
        return 0;
    }
    return sizeof(float) * (uint64_t)(s->nbbox + 4 * s->nvox);
}

struct Collect {
    std::vector<std::shared_ptr<Request>>* batch;
    std::vector<int> owner;  // batch member of each beam of the call
};

int collect_column(void* user, const cccs_sparse_column* c) {
    Collect* k = static_cast<Collect*>(user);
    Request& r = *(*k->batch)[k->owner[c->beam_index]];
    cccs_sparse_column o = *c;
    o.beam_index = 0;
    o.nblocks = 0;  // slot contexts carry no ROI masks
    o.row_block_sizes = nullptr;
    o.keys = reinterpret_cast<const uint64_t*>((uintptr_t)r.keys.size());  // offset, fixed at delivery
    r.keys.insert(r.keys.end(), c->keys, c->keys + c->nnz);
    r.values.insert(r.values.end(), c->values, c->values + c->nnz);
    r.cols.push_back(o);
    return 0;
}

// Per-slot device buffers of the beam-dose requests (fluence map and dose volume).
struct SlotBuffers {
    void* fl = nullptr;
    size_t fl_bytes = 0;
    void* dose = nullptr;
};

int beam_dose_request(cccs_server* s, cccs_ctx* ctx, SlotBuffers& sb, Request& r) {
    const size_t fb = sizeof(float) * r.fluence.size(), db = sizeof(float) * s->nvox;
    if (fb > sb.fl_bytes) {
        if (sb.fl) cccs_dev_free(ctx, sb.fl);
        sb.fl = nullptr;
        sb.fl_bytes = 0;
        CCCS_TRY(cccs_dev_alloc(ctx, fb, &sb.fl));
        sb.fl_bytes = fb;
    }
    if (!sb.dose) CCCS_TRY(cccs_dev_alloc(ctx, db, &sb.dose));
    CCCS_TRY(cccs_dev_upload(ctx, sb.fl, r.fluence.data(), fb));
    CCCS_TRY(cccs_dev_zero(ctx, sb.dose, db));
    CCCS_TRY(cccs_beam_dose_device(ctx, &r.beam, 1, static_cast<const float*>(sb.fl), static_cast<float*>(sb.dose)));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->dose_pool.empty()) {
            r.dose = std::move(s->dose_pool.back());
            s->dose_pool.pop_back();
        }
    }
    r.dose.resize(s->nvox);
    return cccs_dev_download(ctx, r.dose.data(), sb.dose, db);
}

void run_batch(cccs_server* s, cccs_ctx* ctx, SlotBuffers& sb, std::vector<std::shared_ptr<Request>>& batch) {
    if (batch.front()->kind == CCCS_REQ_BEAM_DOSE) {
        for (auto& r : batch) {
            r->status = beam_dose_request(s, ctx, sb, *r);
            if (r->status != CCCS_OK) r->error = cccs_last_error();
        }
        return;
    }
    // beamlet requests: one call over all beams of the batch (shared launches)
    std::vector<cccs_beam> beams;
    std::vector<const float*> fl;
    Collect k{&batch, {}};
    for (size_t i = 0; i < batch.size(); i++) {
        beams.push_back(batch[i]->beam);
        fl.push_back(batch[i]->fluence.data());
        k.owner.push_back((int)i);
    }
    // one call per (threshold, mode) group; in practice the batch shares them
    const float thr = batch.front()->threshold;
    const int mode = batch.front()->mode;
    bool uniform = true;
    for (auto& r : batch) uniform = uniform && r->threshold == thr && r->mode == mode;
    if (uniform) {
        const int st = cccs_beamlet_dose(ctx, beams.data(), (int)beams.size(), fl.data(), thr, mode, collect_column, &k);
        if (st == CCCS_OK) {
            for (auto& r : batch) r->status = CCCS_OK;
            return;
        }
        // one bad request must not fail the others: run them one by one
        for (auto& r : batch) { r->keys.clear(); r->values.clear(); r->cols.clear(); }
    }
    for (size_t i = 0; i < batch.size(); i++) {
        std::vector<std::shared_ptr<Request>> one{batch[i]};
        Collect k1{&one, {0}};
        const float* f1 = batch[i]->fluence.data();
        int st = cccs_beamlet_dose(ctx, &batch[i]->beam, 1, &f1, batch[i]->threshold, batch[i]->mode, collect_column, &k1);
        batch[i]->status = st;
        if (st != CCCS_OK) batch[i]->error = cccs_last_error();
    }
}

void worker(cccs_server* s, int slot) {
    cccs_ctx* ctx = s->slots[slot];
    SlotBuffers sb;
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        uint64_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // the queue this slot may serve now: full-beam requests first; beamlet requests only
            // while another slot is left for full-beam requests
            auto pick = [&]() -> std::deque<std::shared_ptr<Request>>* {
                if (!s->q_dose.empty()) return &s->q_dose;
                if (!s->q_bl.empty() && (s->slots.size() < 2 || s->bl_running < (int)s->slots.size() - 1))
                    return &s->q_bl;
                return nullptr;
            };
            std::deque<std::shared_ptr<Request>>* q = nullptr;
            s->cv_work.wait(lk, [&] {
                q = pick();
                if (q) return s->in_flight == 0 || s->in_flight + q->front()->bytes <= s->budget;
                return s->stop && s->q_dose.empty() && s->q_bl.empty();
            });
            if (!q) {  // stopping, nothing queued
                if (sb.fl) cccs_dev_free(ctx, sb.fl);
                if (sb.dose) cccs_dev_free(ctx, sb.dose);
                return;
            }
            // a full-beam request fills the device by itself and shares no launches with another:
            // it runs alone on its slot so that it completes as soon as its own work does
            const int cap = q == &s->q_dose ? 1 : s->max_batch;
            while (!q->empty() && (int)batch.size() < cap) {
                const auto& h = q->front();
                if (!batch.empty() && s->in_flight + bytes + h->bytes > s->budget) break;
                bytes += h->bytes;
                batch.push_back(h);
                q->pop_front();
            }
            s->in_flight += bytes;
            s->running++;
            if (q == &s->q_bl) s->bl_running++;
        }
        run_batch(s, ctx, sb, batch);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->in_flight -= bytes;
            s->running--;
            if (batch.front()->kind == CCCS_REQ_BEAMLET) s->bl_running--;
            const auto now = Clock::now();
            for (auto& r : batch) {
                r->finished = now;
                r->done = true;
                r->batch_size = (int)batch.size();
                s->latencies.push_back(std::chrono::duration<double, std::milli>(now - r->submitted).count());
            }
            s->batches++;
            s->batched += batch.size();
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace

extern "C" int cccs_server_create(const cccs_problem* problem, int device, int slots, uint64_t hbm_budget,
                                  int max_batch, cccs_server** out) {
    clear_error();
    if (!problem || !out || slots < 1 || slots > 16 || max_batch < 1) return set_error(CCCS_EINVAL, "bad argument");
    *out = nullptr;
    std::unique_ptr<cccs_server> s(new cccs_server());
    for (int i = 0; i < slots; i++) {
        cccs_ctx* c = nullptr;
        int st = cccs_ctx_create(problem, device, &c);
        if (st != CCCS_OK) {
            std::string e = cccs_last_error();
            for (cccs_ctx* d : s->slots) cccs_ctx_destroy(d);
            return set_error(st, "server slot " + std::to_string(i) + ": " + e);
        }
        s->slots.push_back(c);
    }
    s->budget = hbm_budget;
    s->max_batch = max_batch;
    s->nvox = (size_t)problem->size[0] * problem->size[1] * problem->size[2];
    s->nbbox = (size_t)problem->bbox_size[0] * problem->bbox_size[1] * problem->bbox_size[2];
    cccs_server* raw = s.release();
    for (int i = 0; i < slots; i++) raw->workers.emplace_back(worker, raw, i);
    *out = raw;
    return CCCS_OK;
}

extern "C" int cccs_server_submit(cccs_server* s, int kind, const cccs_beam* beam, const float* fluence,
                                  float threshold, int mode, uint64_t* id) {
    clear_error();
    if (!s || !beam || !fluence || !id) return set_error(CCCS_EINVAL, "null argument");
    if (kind != CCCS_REQ_BEAM_DOSE && kind != CCCS_REQ_BEAMLET) return set_error(CCCS_EINVAL, "unknown request kind");
    auto r = std::make_shared<Request>();
    r->kind = kind;
    r->beam = *beam;
    const size_t nf = (size_t)beam->fmap_size[0] * beam->fmap_size[1];
    if (nf == 0 || nf > (size_t)4096 * 4096) return set_error(CCCS_EINVAL, "fluence map dims must be in [1, 4096]");
    r->fluence.assign(fluence, fluence + nf);
    r->threshold = threshold;
    r->mode = mode;
    r->bytes = estimate_bytes(s, s->slots[0], *r);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return set_error(CCCS_EBUSY, "server is shutting down");
        r->id = s->next_id++;
        r->submitted = Clock::now();
        s->requests[r->id] = r;
        (kind == CCCS_REQ_BEAM_DOSE ? s->q_dose : s->q_bl).push_back(r);
        *id = r->id;
    }
    s->cv_work.notify_all();  // the slot rules decide which waiting slot may take it
    return CCCS_OK;
}

extern "C" int cccs_server_wait(cccs_server* s, uint64_t id, float* dose, cccs_sparse_sink sink, void* user,
                                double* latency_ms, int* batch_size) {
    clear_error();
    if (!s) return set_error(CCCS_EINVAL, "null server");
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->requests.find(id);
        if (it == s->requests.end()) return set_error(CCCS_EINVAL, "unknown request id");
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->done; });
        s->requests.erase(it);
    }
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->finished - r->submitted).count();
    if (batch_size) *batch_size = r->batch_size;
    if (r->status != CCCS_OK) return set_error(r->status, r->error);
    if (r->kind == CCCS_REQ_BEAM_DOSE) {
        if (dose) std::memcpy(dose, r->dose.data(), sizeof(float) * r->dose.size());
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->dose_pool.size() < 2 * s->slots.size() + 2) s->dose_pool.push_back(std::move(r->dose));
        return CCCS_OK;
    }
    if (sink) {
        for (cccs_sparse_column c : r->cols) {
            const size_t off = (size_t)(uintptr_t)c.keys;
            c.keys = r->keys.data() + off;
            c.values = r->values.data() + off;
            if (sink(user, &c) != 0) return set_error(CCCS_EIO, "sparse sink aborted");
        }
    }
    return CCCS_OK;
}

extern "C" int cccs_server_preload_angles(cccs_server* s, const cccs_beam* beams, int nbeams, uint64_t max_bytes,
                                          int* nangles) {
    clear_error();
    if (!s || (!beams && nbeams > 0) || nbeams < 0) return set_error(CCCS_EINVAL, "bad argument");
    // the lock is held while the cache is built: submissions wait, and no slot runs meanwhile
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return set_error(CCCS_EBUSY, "server is shutting down");
    if (s->running > 0 || !s->q_dose.empty() || !s->q_bl.empty())
        return set_error(CCCS_EBUSY, "requests in flight; preload before submitting or after they complete");
    for (cccs_ctx* c : s->slots) cccs_ctx_use_angle_cache(c, nullptr);
    cccs_angle_cache_destroy(s->angles);
    s->angles = nullptr;
    if (nangles) *nangles = 0;
    if (nbeams == 0) return CCCS_OK;
    cccs_angle_cache* c = nullptr;
    int st = cccs_angle_cache_create(s->slots[0], beams, nbeams, max_bytes, &c);
    if (st != CCCS_OK) return st;
    for (cccs_ctx* x : s->slots) {
        st = cccs_ctx_use_angle_cache(x, c);
        if (st != CCCS_OK) {
            const std::string e = cccs_last_error();
            for (cccs_ctx* y : s->slots) cccs_ctx_use_angle_cache(y, nullptr);
            cccs_angle_cache_destroy(c);
            return set_error(st, e);
        }
    }
    s->angles = c;
    if (nangles) cccs_angle_cache_info(c, nangles, nullptr);
    return CCCS_OK;
}

extern "C" int cccs_server_stats(cccs_server* s, cccs_server_info* out) {
    clear_error();
    if (!s || !out) return set_error(CCCS_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> v = s->latencies;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
        if (v.empty()) return 0.0;
        const double x = q * (double)(v.size() - 1);
        const size_t i = (size_t)x;
        const double f = x - (double)i;
        return i + 1 < v.size() ? v[i] + f * (v[i + 1] - v[i]) : v[i];
    };
    out->completed = v.size();
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->max_ms = v.empty() ? 0.0 : v.back();
    out->batches = s->batches;
    out->mean_batch = s->batches ? (double)s->batched / (double)s->batches : 0.0;
    out->queued = s->q_dose.size() + s->q_bl.size();
    return CCCS_OK;
}

extern "C" int cccs_server_destroy(cccs_server* s) {
    clear_error();
    if (!s) return CCCS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (std::thread& t : s->workers) t.join();
    for (cccs_ctx* c : s->slots) cccs_ctx_destroy(c);
    cccs_angle_cache_destroy(s->angles);
    delete s;
    return CCCS_OK;
}
