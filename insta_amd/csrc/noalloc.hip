// noalloc.hip -- insta_render_noalloc: insta_render on a fixed device array.
//
// insta_render grows its context's render workspace with hipMalloc when a
// request is larger than any before it.  Callers that must not allocate (a
// loop under HIP graph capture, a serving loop with a fixed memory plan, a
// harness that forbids allocation) call insta_render_noalloc instead: the same
// kernels, the same result, but the round counters and the ray and sample
// buffers are carved from `scratch`, one file-scope device array of
// kNoAllocBytes that each GPU holds from module load to process exit.
//
// Limits: at most kNoAllocMaxPixels pixels per call; each march round holds
// at most kNoAllocSamples samples, so larger frames take further rounds (the
// image is the same; insta_render's workspace usually fits a frame in one).
// Every call on a GPU uses that GPU's one array: calls that overlap on one
// GPU (two host threads, or two streams without ordering) race on it.
#include <mutex>

#include "context.hpp"

namespace insta {
namespace {

constexpr uint32_t kNoAllocMaxPixels = 1u << 16;   // 256 x 256
constexpr uint64_t kNoAllocSamples = 1ull << 20;   // samples per march round
constexpr uint64_t kCounterBytes = 256;            // the round counters, padded
// the workspace carve of render.hip (render_ws_bytes_cap) for the limits above
constexpr uint64_t kNoAllocBytes = kCounterBytes +
                                   (uint64_t)kNoAllocMaxPixels * (3 * 16 + 7 * 4) +
                                   kNoAllocSamples * (4 * 16 + 4) + 15 * 256;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};  // per device: the array's device address

}  // namespace

insta_status noalloc_prepare_device(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("render scratch: device index out of range");
        return INSTA_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return INSTA_OK;
    int cur = -1;
    INSTA_HIP(hipGetDevice(&cur));
    if (cur != device) INSTA_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_ptr[device] = p;
    return INSTA_OK;
}

}  // namespace insta

using insta::set_error;

extern "C" insta_status insta_render_noalloc(insta_ctx* ctx, const insta_camera* cam, uint32_t width, uint32_t height,
                                             int32_t frame, int32_t occupancy, uint32_t spp, float* rgba_dev,
                                             float* depth_dev, uint32_t* steps_dev, void* stream) {
    if (!ctx) {
        set_error("null context");
        return INSTA_ERR_INVALID;
    }
    insta_status st = insta::check_render_args(ctx, cam, width, height, frame, occupancy, rgba_dev);
    if (st != INSTA_OK) return st;
    const uint64_t n_pix = (uint64_t)width * height;
    if (n_pix > insta::kNoAllocMaxPixels) {
        set_error("render_noalloc: at most 65536 pixels (256 x 256) per call");
        return INSTA_ERR_INVALID;
    }
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(insta::g_mu);
        if (ctx->device >= 0 && ctx->device < insta::kMaxDevices) base = insta::g_ptr[ctx->device];
    }
    if (!base) {
        set_error("render_noalloc: the device's render scratch was not resolved at context creation");
        return INSTA_ERR_STATE;
    }
    static_assert(insta::kRenderCounterWords * sizeof(uint32_t) <= insta::kCounterBytes, "counter words");
    if (insta::kCounterBytes + insta::render_ws_bytes_cap((uint32_t)n_pix, insta::kNoAllocSamples) >
        insta::kNoAllocBytes) {
        set_error("render_noalloc: workspace carve exceeds the scratch array");
        return INSTA_ERR_INVALID;
    }
    insta::RenderWS ws;
    uint32_t h_counters[insta::kRenderCounterWords];
    ws.d_counters = static_cast<uint32_t*>(base);
    ws.h_counters = h_counters;
    insta::bind_render_ws_cap(ws, static_cast<char*>(base) + insta::kCounterBytes, (uint32_t)n_pix,
                              insta::kNoAllocSamples);
    st = insta::render(ctx, ws, cam, width, height, frame, occupancy, spp, rgba_dev, depth_dev, steps_dev,
                       (hipStream_t)stream);
    if (st == INSTA_OK) ctx->ws.stats = ws.stats;  // insta_get_render_stats reports the last render
    return st;
}
