// hana_server.cpp -- continuous-batching render server (hana_server_* in
// include/hana_gpu.h).
//
// A worker thread owns `lanes` internal contexts, each with its own HIP
// stream and workspace.  Whenever a lane is idle it takes the oldest pending
// requests that fit the batch limits and the HBM budget, renders them as one
// launch sequence (every request is a DrawModel::draw frame on cleared
// targets, SR/scene.h:53-99), copies each colour image to pinned host memory
// and records an event.  Lanes run concurrently, so requests that arrive while
// a batch is on the GPU start on another lane as soon as one is free.  Each
// request's passes only read and write its own targets, so its image is the
// same whatever batch it runs in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "hana_hostmath.h"
#include "hana_internal.h"

using namespace hana;
using namespace hana::rt;

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t id = 0;
    hana_request r{};
    hana_uniforms u{};
    double t_submit = 0, t_launch = 0;
    int64_t bytes = 0;            // admitted device footprint
    int sw = 0, sh = 0;           // shadow map size
    hana_target* target = nullptr;
    hana_target* shadow = nullptr;
    uint8_t* host = nullptr;
    size_t host_cap = 0;
};

struct Lane {
    hana_ctx* ctx = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    int retries = 0;
    std::vector<Req> batch;
};

int64_t px_bucket(int64_t px) {
    int64_t b = 1 << 16;
    while (b < px) b <<= 1;
    return b;
}

}  // namespace

struct hana_server {
    int device = 0;
    hana_server_config cfg{};
    std::vector<Lane> lanes;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    std::deque<Req> pending;
    std::deque<hana_completion> done;
    std::map<uint64_t, std::pair<uint8_t*, size_t>> held;     // id -> pinned image
    std::multimap<size_t, uint8_t*> pinned_pool;              // capacity -> buffer
    std::multimap<int64_t, hana_target*> target_pool;         // worker thread only
    std::atomic<bool> stop{false};
    uint64_t next_id = 1;
    int64_t inflight_bytes = 0;
    hana_server_stats st{};
    Clock::time_point t0 = Clock::now();
    double busy_since = -1;
    std::string err;

    double now() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

namespace {

// Device footprint charged to a request: its targets and the per-pass
// workspace (per vertex: position, varyings, outcode; per face: bins, clip
// queue and pool, triangle list).
int64_t request_bytes(const hana_request& r, int sw, int sh) {
    int64_t V = r.mesh->V, F = r.mesh->F;
    int passes = r.enable_shadow ? 2 : 1;
    int64_t b = px_bucket((int64_t)r.width * r.height) * 8;
    if (r.enable_shadow) b += px_bucket((int64_t)sw * sh) * 8;
    b += passes * (V * 56 + F * 84);
    return b;
}

hana_target* get_target(hana_server* s, int w, int h) {
    int64_t cap = px_bucket((int64_t)w * h);
    auto it = s->target_pool.find(cap);
    hana_target* t = nullptr;
    if (it != s->target_pool.end()) {
        t = it->second;
        s->target_pool.erase(it);
    } else {
        t = new hana_target();
        if (hipMalloc((void**)&t->color, 4 * cap) != hipSuccess || hipMalloc((void**)&t->depth, 4 * cap) != hipSuccess) {
            (void)hipGetLastError();
            if (t->color) (void)hipFree(t->color);
            delete t;
            return nullptr;
        }
    }
    t->w = w;
    t->h = h;
    t->pending_clear = true;          // main.cpp's per-frame clear (Black, FLT_MAX)
    t->clear_rgba = 0xFF000000u;
    t->clear_depth = 3.40282347e+38f;
    return t;
}

void put_target(hana_server* s, hana_target* t) {
    if (t) s->target_pool.insert({px_bucket((int64_t)t->w * t->h), t});
}

uint8_t* get_pinned(hana_server* s, size_t bytes, size_t* cap) {
    size_t c = (size_t)px_bucket((int64_t)((bytes + 3) / 4)) * 4;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->pinned_pool.find(c);
        if (it != s->pinned_pool.end()) {
            uint8_t* p = it->second;
            s->pinned_pool.erase(it);
            *cap = c;
            return p;
        }
    }
    uint8_t* p = nullptr;
    if (hipHostMalloc((void**)&p, c) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    *cap = c;
    return p;
}

// Enqueues one batch on a lane.
int launch(hana_server* s, Lane& L) {
    std::vector<PassSpec> specs;
    for (Req& q : L.batch) {
        if (!q.target) q.target = get_target(s, q.r.width, q.r.height);
        if (q.r.enable_shadow && !q.shadow) q.shadow = get_target(s, q.sw, q.sh);
        size_t img = (size_t)q.r.width * q.r.height * 4;
        if (!q.host) q.host = get_pinned(s, (s->cfg.flags & HANA_SERVER_RETURN_DEPTH) ? 2 * img : img, &q.host_cap);
        if (!q.target || (q.r.enable_shadow && !q.shadow) || !q.host) return HANA_E_NOMEM;
        q.target->pending_clear = true;
        if (q.shadow) q.shadow->pending_clear = true;
        hana_uniforms u = q.u;
        u.enable_shadow = q.r.enable_shadow;
        if (q.r.enable_shadow)
            specs.push_back({q.r.mesh, HANA_SHADER_SHADOW, u, q.r.diffuse, q.r.normal, nullptr, q.shadow, 0});
        specs.push_back({q.r.mesh, q.r.shader, u, q.r.diffuse, q.r.normal, q.r.enable_shadow ? q.shadow : nullptr,
                         q.target, 1});
    }
    int rc = run_passes(L.ctx, std::move(specs));
    if (rc) return rc;
    for (Req& q : L.batch) {
        size_t img = (size_t)q.r.width * q.r.height * 4;
        if (hipMemcpyAsync(q.host, q.target->color, img, hipMemcpyDeviceToHost, L.ctx->stream) != hipSuccess)
            return HANA_E_DEVICE;
        if ((s->cfg.flags & HANA_SERVER_RETURN_DEPTH) &&
            hipMemcpyAsync(q.host + img, q.target->depth, img, hipMemcpyDeviceToHost, L.ctx->stream) != hipSuccess)
            return HANA_E_DEVICE;
    }
    if (hipEventRecord(L.done, L.ctx->stream) != hipSuccess) return HANA_E_DEVICE;
    L.busy = true;
    return HANA_OK;
}

void complete(hana_server* s, Lane& L, int status) {
    double t = s->now();
    std::lock_guard<std::mutex> g(s->mu);
    for (Req& q : L.batch) {
        hana_completion c{};
        c.id = q.id;
        c.user_tag = q.r.user_tag;
        c.status = status;
        c.width = q.r.width;
        c.height = q.r.height;
        c.batch_frames = (int)L.batch.size();
        c.rgba = status == HANA_OK ? q.host : nullptr;
        c.depth = status == HANA_OK && (s->cfg.flags & HANA_SERVER_RETURN_DEPTH)
                      ? reinterpret_cast<const float*>(q.host + (size_t)q.r.width * q.r.height * 4)
                      : nullptr;
        c.t_submit_ms = q.t_submit;
        c.t_launch_ms = q.t_launch;
        c.t_done_ms = t;
        s->done.push_back(c);
        if (status == HANA_OK) s->held[q.id] = {q.host, q.host_cap};
        else if (q.host) s->pinned_pool.insert({q.host_cap, q.host});
        put_target(s, q.target);
        put_target(s, q.shadow);
        s->inflight_bytes -= q.bytes;
        s->st.completed += 1;
    }
    L.batch.clear();
    L.busy = false;
    L.retries = 0;
    s->cv_done.notify_all();
}

// Takes the oldest pending requests that fit the limits and the budget.
std::vector<Req> take_batch(hana_server* s) {
    std::vector<Req> b;
    std::lock_guard<std::mutex> g(s->mu);
    int64_t px = 0;
    while (!s->pending.empty() && (int)b.size() < s->cfg.max_batch_frames) {
        Req& q = s->pending.front();
        int64_t qpx = (int64_t)q.r.width * q.r.height + (q.r.enable_shadow ? (int64_t)q.sw * q.sh : 0);
        if (!b.empty() && px + qpx > s->cfg.max_batch_pixels) break;
        if (q.bytes > s->cfg.hbm_budget_bytes) {   // can never fit: fail it
            hana_completion c{};
            c.id = q.id;
            c.user_tag = q.r.user_tag;
            c.status = HANA_E_NOMEM;
            c.t_submit_ms = c.t_launch_ms = c.t_done_ms = s->now();
            s->done.push_back(c);
            s->st.rejected += 1;
            s->pending.pop_front();
            s->cv_done.notify_all();
            continue;
        }
        if (s->inflight_bytes + q.bytes > s->cfg.hbm_budget_bytes) break;   // wait for memory to free up
        s->inflight_bytes += q.bytes;
        s->st.peak_inflight_bytes = std::max(s->st.peak_inflight_bytes, s->inflight_bytes);
        px += qpx;
        b.push_back(q);
        s->pending.pop_front();
    }
    return b;
}

void worker_main(hana_server* s) {
    (void)hipSetDevice(s->device);
    while (true) {
        bool any_busy = false;
        for (Lane& L : s->lanes) {
            if (!L.busy) continue;
            hipError_t q = hipEventQuery(L.done);
            if (q == hipErrorNotReady) {
                any_busy = true;
                continue;
            }
            int rc = (q == hipSuccess) ? check_status(L.ctx) : HANA_E_DEVICE;
            if (rc == HANA_E_OVERFLOW && L.retries < 3) {
                // workspace grown by check_status: render the same batch again
                L.retries++;
                s->st.retries++;
                if (launch(s, L) == HANA_OK) {
                    any_busy = true;
                    continue;
                }
                rc = HANA_E_DEVICE;
            }
            complete(s, L, rc);
        }
        if (s->stop.load() && !any_busy) {
            std::lock_guard<std::mutex> g(s->mu);
            if (s->pending.empty()) break;
        }
        for (Lane& L : s->lanes) {
            if (L.busy) continue;
            std::vector<Req> b = take_batch(s);
            if (b.empty()) break;
            double t = s->now();
            for (Req& q : b) q.t_launch = t;
            L.batch = std::move(b);
            int rc = launch(s, L);
            if (rc) {
                complete(s, L, rc);
                continue;
            }
            s->st.batches++;
            any_busy = true;
        }
        double t = s->now();
        if (any_busy) {
            if (s->busy_since < 0) s->busy_since = t;
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_submit.wait_for(lk, std::chrono::microseconds(50));
        } else {
            if (s->busy_since >= 0) {
                s->st.busy_ms += t - s->busy_since;
                s->busy_since = -1;
            }
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_submit.wait_for(lk, std::chrono::milliseconds(2),
                                  [&] { return !s->pending.empty() || s->stop.load(); });
        }
    }
}

}  // namespace

extern "C" {

int hana_server_create(hana_ctx* ctx, const hana_server_config* cfg, hana_server** out) {
    if (!ctx || !cfg || !out || cfg->lanes < 1 || cfg->lanes > 8 || cfg->max_batch_frames < 1 ||
        cfg->max_batch_pixels < 1 || cfg->hbm_budget_bytes < 1 || (cfg->flags & ~HANA_SERVER_RETURN_DEPTH))
        return fail(ctx, HANA_E_INVALID, "hana_server_create: bad configuration");
    hana_server* s = new hana_server();
    s->device = ctx->device;
    s->cfg = *cfg;
    for (int i = 0; i < cfg->lanes; i++) {
        Lane L;
        // Lane contexts hold three streams though a lane draws on one: lanes
        // of four-stream contexts served C4 with p99 2.5 ms, three-stream ones
        // with 1.0 ms; likely because streams map round-robin onto the four
        // hardware queues of a process, so the lanes' own streams then share one.
        if (create_context(ctx->device, 3, &L.ctx) != HANA_OK ||
            hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess) {
            if (L.ctx) hana_gpu_destroy(L.ctx);
            for (Lane& l : s->lanes) {
                hana_gpu_destroy(l.ctx);
                (void)hipEventDestroy(l.done);
            }
            delete s;
            return fail(ctx, HANA_E_DEVICE, "hana_server_create: lane set-up failed");
        }
        // a lane's completion event is recorded on its context stream, so its
        // draws stay on that one stream; a batch that outgrew the workspace is
        // launched again by the worker (with its image copies)
        L.ctx->pipelined = false;
        L.ctx->replay_overflow = false;
        s->lanes.push_back(L);
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return HANA_OK;
}

int hana_server_submit(hana_server* s, const hana_request* r, uint64_t* id) {
    if (!s || !r || !r->mesh || !valid_kind(r->shader) || r->shader == HANA_SHADER_SHADOW || r->width <= 0 ||
        r->height <= 0 || (int64_t)r->width * r->height > (1ll << 26))
        return HANA_E_INVALID;
    Req q;
    q.r = *r;
    q.sw = r->shadow_width > 0 ? r->shadow_width : r->width;
    q.sh = r->shadow_height > 0 ? r->shadow_height : r->height;
    const float zero[3] = {0, 0, 0}, one[3] = {1, 1, 1};
    float aspect = (float)r->width / (float)r->height;
    build_uniforms(r->cam_pos, r->cam_target, aspect, r->light_pos, zero, zero, one, r->width, r->height,
                   r->enable_shadow, &q.u);
    q.u.shadow_fp32 = r->shadow_fp32 ? 1 : 0;
    q.bytes = request_bytes(*r, q.sw, q.sh);
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop.load()) return HANA_E_INVALID;
        q.id = s->next_id++;
        q.t_submit = s->now();
        if (id) *id = q.id;
        s->pending.push_back(q);
        s->st.submitted++;
    }
    s->cv_submit.notify_one();
    return HANA_OK;
}

int hana_server_poll(hana_server* s, hana_completion* out, int max, int* n, int wait_ms) {
    if (!s || !out || !n || max < 1) return HANA_E_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
    int k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *n = k;
    return HANA_OK;
}

int hana_server_release(hana_server* s, uint64_t id) {
    if (!s) return HANA_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->held.find(id);
    if (it == s->held.end()) return HANA_E_INVALID;
    s->pinned_pool.insert({it->second.second, it->second.first});
    s->held.erase(it);
    return HANA_OK;
}

int hana_server_get_stats(hana_server* s, hana_server_stats* out) {
    if (!s || !out) return HANA_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->st;
    out->inflight_bytes = s->inflight_bytes;
    if (s->busy_since >= 0) out->busy_ms += s->now() - s->busy_since;
    return HANA_OK;
}

double hana_server_now_ms(hana_server* s) { return s ? s->now() : 0.0; }

int hana_server_destroy(hana_server* s) {
    if (!s) return HANA_E_INVALID;
    s->stop.store(true);
    s->cv_submit.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (Lane& L : s->lanes) {
        (void)hipStreamSynchronize(L.ctx->stream);
        (void)hipEventDestroy(L.done);
        hana_gpu_destroy(L.ctx);
    }
    for (auto& kv : s->target_pool) {
        (void)hipFree(kv.second->color);
        (void)hipFree(kv.second->depth);
        delete kv.second;
    }
    for (auto& kv : s->pinned_pool) (void)hipHostFree(kv.second);
    for (auto& kv : s->held) (void)hipHostFree(kv.second.first);
    delete s;
    return HANA_OK;
}

}  // extern "C"
