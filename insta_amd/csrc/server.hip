// server.hip -- request loop over several resident avatars on one GPU.
//
// The reference has no server: it renders one frame at a time from
// Testbed::render_frame (instant-ngp/src/testbed.cu:2613) with one stream.
// Here a request (avatar, 50 expression coefficients, camera, resolution)
// runs the per-request preparation -- expression blend into the avatar's
// frame slot, device BVH and per-triangle frames (insta_set_frame_expression),
// occupancy grid of the posed mesh -- then the render pipeline.
//
// Continuous batching: `n_slots` workers, each a host thread with its own
// HIP stream (default priority), frame slot and render workspace.  Requests
// queue FIFO; the head request is admitted as soon as a worker is idle and
// its render workspace fits the HBM budget, carved from one device arena
// (first fit, no allocation per request).  Admitted requests overlap on the
// device: their kernels run on different streams, while each worker's host
// thread waits only on its own stream.  Each request's result depends only on
// its own inputs (no shared scratch between slots), so it is the same alone
// and in a mix.  Each request's whole chain -- preparation, ray init, two
// march/deform/network/composite rounds with live counts kept on the device,
// image write -- is stream-captured and moved onto the worker's executable
// graph for that avatar with hipGraphExecUpdate (same shape, new arguments:
// camera, resolution, workspace offset, output), then launched as one graph;
// the host synchronises once per request, and only rays still live after the
// two rounds take further host-checked rounds.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "context.hpp"

namespace insta {
namespace {

constexpr uint64_t kArenaAlign = 256;
inline uint64_t arena_bytes(uint64_t bytes) { return (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign; }

// first-fit allocator over [0, size) with coalescing frees
class Arena {
public:
    void reset(uint64_t size) {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(uint64_t bytes, uint64_t& off) {
        bytes = arena_bytes(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const uint64_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = arena_bytes(bytes);
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<uint64_t, uint64_t> free_;
};

struct Pending {
    uint64_t ticket;
    insta_request req;
    std::vector<float> coeffs;
    double t_submit;
    uint64_t bytes;
};

}  // namespace
}  // namespace insta

struct insta_server {
    struct Avatar {
        insta_ctx* ctx;
        const float* bases;
        uint32_t n_bases;
    };
    struct Slot {
        uint32_t index = 0;
        hipStream_t stream = nullptr;
        insta::RenderWS ws;
        float* d_coeffs = nullptr;
        float* h_coeffs = nullptr;
        std::vector<hipGraphExec_t> graphs;  // per avatar: preparation + render of the last request
        uint64_t graph_updates = 0, graph_builds = 0;
        std::thread thread;
    };

    insta_server_config cfg{};
    int device = 0;
    uint32_t max_bases = 0;
    char* arena_base = nullptr;
    insta::Arena arena;
    std::vector<Avatar> avatars;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<insta::Pending> queue;
    std::vector<insta_result> done;
    uint64_t next_ticket = 1;
    uint32_t running = 0;
    bool stop = false;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace insta {
namespace {

constexpr uint32_t kMaxServerBases = 256;

// expression -> BVH/frames -> occupancy of the worker's frame slot
insta_status prepare(insta_server* S, insta_server::Slot& sl, const insta_server::Avatar& av, hipStream_t stream) {
    insta_ctx* ctx = av.ctx;
    const uint32_t slot = sl.index + 1;
    insta_status st = launch_blend_expression(ctx, slot, av.bases, av.n_bases, sl.d_coeffs, stream);
    if (st != INSTA_OK) return st;
    st = launch_extra_from_coeffs(ctx, sl.index, sl.d_coeffs, av.n_bases, stream);  // no-op without extra dims
    if (st != INSTA_OK) return st;
    st = launch_set_built(ctx, slot, 1, stream);
    if (st != INSTA_OK) return st;
    return launch_mesh_occupancy(ctx, slot, S->cfg.occupancy_threshold, stream);
}

// Render rounds queued with each request (live counts stay on the device);
// rays still live after them continue on host-checked rounds.
constexpr uint32_t kQueuedRounds = 2;

insta_status run_request(insta_server* S, insta_server::Slot& sl, Pending& p, char* ws_base) {
    const insta_request& r = p.req;
    const insta_server::Avatar& av = S->avatars[r.avatar];
    insta_ctx* ctx = av.ctx;
    hipStream_t stream = sl.stream;
    std::copy(p.coeffs.begin(), p.coeffs.end(), sl.h_coeffs);
    INSTA_HIP(hipMemcpyAsync(sl.d_coeffs, sl.h_coeffs, av.n_bases * sizeof(float), hipMemcpyHostToDevice, stream));
    bind_render_ws(sl.ws, ws_base, r.width * r.height, S->cfg.samples_per_pixel);
    const int32_t frame = (int32_t)sl.index;
    auto enqueue = [&]() -> insta_status {
        insta_status st = prepare(S, sl, av, stream);
        if (st != INSTA_OK) return st;
        return enqueue_render(ctx, sl.ws, &r.camera, r.width, r.height, frame, frame, 0, kQueuedRounds, r.rgba_dev,
                              nullptr, nullptr, stream);
    };
    insta_status st;
    if (S->cfg.use_graphs) {
        // capture this request's chain and move the slot's executable graph
        // for this avatar onto it (same topology: only arguments change)
        hipGraph_t graph = nullptr;
        INSTA_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        st = enqueue();
        const hipError_t ce = hipStreamEndCapture(stream, &graph);
        if (st != INSTA_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        INSTA_HIP(ce);
        hipGraphExec_t& exec = sl.graphs[r.avatar];
        bool updated = false;
        if (exec) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            updated = hipGraphExecUpdate(exec, graph, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(exec);
                exec = nullptr;
            }
        }
        hipError_t ie = hipSuccess;
        if (!updated) ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        INSTA_HIP(ie);
        ++(updated ? sl.graph_updates : sl.graph_builds);
        INSTA_HIP(hipGraphLaunch(exec, stream));
    } else {
        st = enqueue();
        if (st != INSTA_OK) return st;
    }
    // host-side state the queued kernels established
    ctx->slot_built[sl.index + 1] = 1;
    ctx->grid_built[sl.index + 1] = 1;
    st = finish_render(ctx, sl.ws, &r.camera, r.width, r.height, frame, frame, 0, kQueuedRounds, r.rgba_dev, nullptr,
                       nullptr, stream);
    if (st != INSTA_OK) return st;
    INSTA_HIP(hipStreamSynchronize(stream));
    return INSTA_OK;
}

void worker(insta_server* S, insta_server::Slot* sl) {
    (void)hipSetDevice(S->device);
    for (;;) {
        Pending p;
        uint64_t off = 0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            // FIFO admission: the head request, once its workspace fits
            S->cv.wait(lk, [&] {
                return (S->stop && S->queue.empty()) || (!S->queue.empty() && S->arena.alloc(S->queue.front().bytes, off));
            });
            if (S->queue.empty()) return;
            p = std::move(S->queue.front());
            S->queue.pop_front();
            ++S->running;
        }
        insta_result res{};
        res.ticket = p.ticket;
        res.user_tag = p.req.user_tag;
        res.t_submit = p.t_submit;
        res.t_admit = S->now();
        res.slot = sl->index;
        const insta_status st = run_request(S, *sl, p, S->arena_base + off);
        res.status = st;
        res.n_samples = sl->ws.stats.n_samples;
        res.t_done = S->now();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->arena.release(off, p.bytes);
            --S->running;
            S->done.push_back(res);
        }
        S->cv.notify_all();
    }
}

}  // namespace
}  // namespace insta

extern "C" {

void insta_default_server_config(insta_server_config* c) {
    if (!c) return;
    c->n_slots = 6;
    c->hbm_budget_bytes = 24ull << 30;
    c->samples_per_pixel = 64;
    c->occupancy_threshold = 0.05f;
    c->use_graphs = 1;
}

insta_status insta_server_create(const insta_server_config* cfg, insta_server** out) {
    using insta::set_error;
    if (!cfg || !out || cfg->n_slots == 0 || cfg->n_slots > 64 || cfg->samples_per_pixel == 0 ||
        cfg->hbm_budget_bytes < (64ull << 20) || !(cfg->occupancy_threshold > 0.0f)) {
        set_error("server_create: bad configuration");
        return INSTA_ERR_INVALID;
    }
    *out = nullptr;
    insta_server* S = new (std::nothrow) insta_server();
    if (!S) {
        set_error("host allocation failed");
        return INSTA_ERR_OOM;
    }
    S->cfg = *cfg;
    // the arena hands out 256-byte multiples: a budget that is not one could
    // accept a request (in submit) that never fits and then blocks the queue
    S->cfg.hbm_budget_bytes = cfg->hbm_budget_bytes / insta::kArenaAlign * insta::kArenaAlign;
    INSTA_HIP(hipGetDevice(&S->device));
    if (hipMalloc(reinterpret_cast<void**>(&S->arena_base), S->cfg.hbm_budget_bytes) != hipSuccess) {
        delete S;
        set_error("server_create: cannot reserve the HBM budget");
        return INSTA_ERR_OOM;
    }
    S->arena.reset(S->cfg.hbm_budget_bytes);
    for (uint32_t i = 0; i < cfg->n_slots; ++i) {
        auto sl = std::make_unique<insta_server::Slot>();
        sl->index = i;
        bool ok = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) == hipSuccess &&
                  hipMalloc(reinterpret_cast<void**>(&sl->d_coeffs), insta::kMaxServerBases * sizeof(float)) ==
                      hipSuccess &&
                  hipHostMalloc(reinterpret_cast<void**>(&sl->h_coeffs), insta::kMaxServerBases * sizeof(float),
                                hipHostMallocDefault) == hipSuccess &&
                  hipMalloc(reinterpret_cast<void**>(&sl->ws.d_counters), 16 * sizeof(uint32_t)) == hipSuccess &&
                  hipHostMalloc(reinterpret_cast<void**>(&sl->ws.h_counters), 16 * sizeof(uint32_t),
                                hipHostMallocDefault) == hipSuccess;
        S->slots.push_back(std::move(sl));
        if (!ok) {
            insta_server_destroy(S);
            set_error("server_create: slot allocation failed");
            return INSTA_ERR_OOM;
        }
    }
    for (auto& sl : S->slots) sl->thread = std::thread(insta::worker, S, sl.get());
    *out = S;
    return INSTA_OK;
}

insta_status insta_server_add_avatar(insta_server* S, insta_ctx* ctx, const float* bases_dev, uint32_t n_bases,
                                     uint32_t* avatar_id) {
    using insta::set_error;
    if (!S || !ctx || !bases_dev || !avatar_id || n_bases == 0 || n_bases > insta::kMaxServerBases) {
        set_error("server_add_avatar: bad argument");
        return INSTA_ERR_INVALID;
    }
    if (!ctx->params_set || !ctx->slot_built[0]) {
        set_error("server_add_avatar: the avatar needs its canonical mesh and parameters");
        return INSTA_ERR_STATE;
    }
    if (ctx->cfg.max_frames < S->cfg.n_slots) {
        set_error("server_add_avatar: max_frames must be >= n_slots (one frame slot per worker)");
        return INSTA_ERR_INVALID;
    }
    if (!insta::mfma_network_supported(ctx->cfg) &&
        (!insta::network_supported(ctx->cfg) || !insta::scalar_precision(ctx->cfg))) {
        set_error("server_add_avatar: configuration not supported by the fused MLP kernels");
        return INSTA_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->next_ticket != 1) {
        set_error("server_add_avatar: avatars must be added before the first request");
        return INSTA_ERR_STATE;
    }
    S->avatars.push_back({ctx, bases_dev, n_bases});
    for (auto& sl : S->slots) sl->graphs.push_back(nullptr);
    *avatar_id = (uint32_t)S->avatars.size() - 1;
    return INSTA_OK;
}

insta_status insta_server_submit(insta_server* S, const insta_request* r, uint64_t* ticket) {
    using insta::set_error;
    if (!S || !r || !r->coeffs || !r->rgba_dev || r->width == 0 || r->height == 0 ||
        (uint64_t)r->width * r->height > (1ull << 24)) {
        set_error("server_submit: bad request");
        return INSTA_ERR_INVALID;
    }
    if (r->avatar >= S->avatars.size()) {
        set_error("server_submit: unknown avatar");
        return INSTA_ERR_INVALID;
    }
    insta::Pending p;
    p.req = *r;
    p.coeffs.assign(r->coeffs, r->coeffs + S->avatars[r->avatar].n_bases);
    p.bytes = insta::render_ws_bytes(r->width * r->height, S->cfg.samples_per_pixel);
    // compared as the arena will carve it (rounded up to its alignment)
    if (insta::arena_bytes(p.bytes) > S->cfg.hbm_budget_bytes) {
        set_error("server_submit: the request's workspace exceeds the HBM budget");
        return INSTA_ERR_INVALID;
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        p.ticket = S->next_ticket++;
        p.t_submit = S->now();
        if (ticket) *ticket = p.ticket;
        S->queue.push_back(std::move(p));
    }
    S->cv.notify_all();
    return INSTA_OK;
}

insta_status insta_server_poll(insta_server* S, insta_result* out, uint32_t max_results, uint32_t* n_out, int wait) {
    if (!S || !n_out || (max_results && !out)) {
        insta::set_error("server_poll: bad argument");
        return INSTA_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    if (wait) S->cv.wait(lk, [&] { return !S->done.empty() || (S->queue.empty() && S->running == 0); });
    const uint32_t n = std::min<uint32_t>(max_results, (uint32_t)S->done.size());
    std::copy(S->done.begin(), S->done.begin() + n, out);
    S->done.erase(S->done.begin(), S->done.begin() + n);
    *n_out = n;
    return INSTA_OK;
}

double insta_server_clock(insta_server* S) { return S ? S->now() : 0.0; }

insta_status insta_server_graph_stats(insta_server* S, uint64_t* updates, uint64_t* builds) {
    if (!S || !updates || !builds) {
        insta::set_error("server_graph_stats: bad argument");
        return INSTA_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    *updates = *builds = 0;
    for (auto& sl : S->slots) {
        *updates += sl->graph_updates;
        *builds += sl->graph_builds;
    }
    return INSTA_OK;
}

void insta_server_destroy(insta_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    for (auto& sl : S->slots)
        if (sl->thread.joinable()) sl->thread.join();
    for (auto& sl : S->slots) {
        for (hipGraphExec_t e : sl->graphs)
            if (e) (void)hipGraphExecDestroy(e);
        insta::release_render_ws(sl->ws);
        insta::dev_free(sl->d_coeffs);
        if (sl->h_coeffs) (void)hipHostFree(sl->h_coeffs);
        if (sl->stream) (void)hipStreamDestroy(sl->stream);
    }
    insta::dev_free(S->arena_base);
    delete S;
}

}  // extern "C"
