// render.hip -- ray init, occupancy-grid march and front-to-back composite.
//
// Restates NerfTracer::init_rays_from_camera + trace (instant-ngp/src/
// testbed_nerf.cu:2030-2244) for a single cascade and cone angle 0 (the
// reference's unit-cube case, testbed_nerf.cu:2684-2691), with this schedule:
//   init      one thread per pixel: pinhole ray, render-box entry, one-step
//             ld jitter and empty-space skip; live rays appended to a list
//   repeat    march: each live ray emits up to K samples (K = sample
//                    capacity / live rays, computed on the device), one
//                    walk with t to scratch; a wave's samples take one
//                    contiguous range (one atomic), step-major inside it:
//                    step j of the wave's 64 adjacent rays side by side, so
//                    the deform and network see neighbouring points in
//                    neighbouring lanes (the reference's sample index is
//                    likewise step-major, i + j * n_alive)
//             deform: posed -> canonical per sample (deform.hip)
//             network: encode + MLPs per sample (network.hip)
//             composite: per ray over its samples; rays that neither hit
//                    the transmittance threshold, left the box nor reached
//                    max_steps are appended to the next live list
//   finalize  one thread per pixel writes RGBA/depth/step counts.
// Each ray's samples depend only on its own t sequence, so the result does
// not depend on K or on the order rays are processed in.  Live counts stay on
// the device: the first round is queued with init and finalize, and the host
// reads the counters once, adding rounds only for rays still live.
#include <cmath>
#include <type_traits>

#include "context.hpp"
#include "march.hpp"

namespace insta {
namespace {

constexpr uint32_t kRayThreads = 256;

enum : uint32_t {
    kCntAliveA = 0,
    kCntAliveB = 1,
    kCntSamples = 2,
    kCntAlive0 = 3,
    kCntComposited = 4,
    kCntEmitted = 5,
    kCntRounds = 6,
    kCntMisses = 7,   // in-box samples whose nearest-triangle query found nothing
    kNumCounters = 16
};

// Threads walk the image in 8x8 pixel tiles (one wave per tile), so the live
// list -- and with it the samples of a march wave -- holds square patches of
// neighbouring rays: their samples sit close together, the nearest-triangle
// searches of a wave follow nearly the same BVH path and the hash-grid
// gathers share cache lines.  Ray state stays indexed by pixel.
constexpr uint32_t kTile = 8;
INSTA_HD uint32_t tiled_threads(uint32_t w, uint32_t h) {
    return ((w + kTile - 1) / kTile) * ((h + kTile - 1) / kTile) * kTile * kTile;
}

__global__ __launch_bounds__(kRayThreads) void init_rays_kernel(RenderArgs a, float4* __restrict__ r_o,
                                                                float4* __restrict__ r_d, float4* __restrict__ r_rgba,
                                                                float* __restrict__ r_depth,
                                                                uint32_t* __restrict__ r_steps,
                                                                uint32_t* __restrict__ alive,
                                                                uint32_t* __restrict__ counters) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t tiles_x = (a.width + kTile - 1) / kTile;
    const uint32_t tile = tid / (kTile * kTile), in_tile = tid % (kTile * kTile);
    const uint32_t x = (tile % tiles_x) * kTile + in_tile % kTile;
    const uint32_t y = (tile / tiles_x) * kTile + in_tile / kTile;
    const uint32_t i = y * a.width + x;  // pixel index
    bool live = false;
    if (x < a.width && y < a.height) {
        // pixel_to_ray, pinhole (common_device.cuh:260-315)
        const float u = ((float)x + a.off_x) / (float)a.width;
        const float v = ((float)y + a.off_y) / (float)a.height;
        const V3 d0 = v3((u - a.cx) * (float)a.width / a.fx, (v - a.cy) * (float)a.height / a.fy, 1.0f);
        V3 d = v3(a.M[0] * d0.x + a.M[1] * d0.y + a.M[2] * d0.z, a.M[4] * d0.x + a.M[5] * d0.y + a.M[6] * d0.z,
                  a.M[8] * d0.x + a.M[9] * d0.y + a.M[10] * d0.z);
        const float dn = norm(d);
        d = v3(d.x / dn, d.y / dn, d.z / dn);
        const V3 o = v3(a.M[3], a.M[7], a.M[11]);
        float t = fmaxf(box_ray_tmin(a, o, d), 0.05f) + 1e-6f;
        if (box_contains(a, o + d * t)) {
            const V3 idir = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
            t += ld_random_val0(a.spp, i * 786433u) * a.dt;  // advance_pos_nerf (testbed_nerf.cu:663-664)
            V3 pos;
            live = skip_empty(a, o, d, idir, t, pos);
        }
        r_o[i] = make_float4(o.x, o.y, o.z, t);
        r_d[i] = make_float4(d.x, d.y, d.z, 0.0f);
        r_rgba[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        r_depth[i] = 0.0f;
        r_steps[i] = 0;
    }
    const uint32_t slot = wave_append(&counters[kCntAliveA], live);
    if (live) alive[slot] = i;
    // counters[kCntAlive0] is derived from the first live count on the host
}

// One walk per ray: the t of each emitted sample goes to step-major scratch
// (t_buf[j * n_alive + k], coalesced across the wave; K * n_alive fits the
// sample buffer), then the wave reserves contiguous output space and each
// lane writes its samples from the stored t values (pos = o + d t exactly as
// during the walk).
__global__ __launch_bounds__(kRayThreads) void march_kernel(RenderArgs a, const uint32_t* __restrict__ alive,
                                                            const uint32_t* __restrict__ n_alive_dev,
                                                            uint64_t sample_cap, float4* __restrict__ r_o,
                                                            const float4* __restrict__ r_d,
                                                            const uint32_t* __restrict__ r_steps,
                                                            uint32_t* __restrict__ r_flags,
                                                            uint32_t* __restrict__ r_off, uint32_t* __restrict__ r_cnt,
                                                            float* __restrict__ t_buf, float4* __restrict__ s_pos,
                                                            uint32_t* __restrict__ counters) {
    const uint32_t n_alive = *n_alive_dev;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if ((k & ~63u) >= n_alive) return;  // whole wave past the live list
    // samples per live ray this round: all fit the sample buffer
    uint64_t K64 = sample_cap / n_alive;
    if (K64 > a.max_steps) K64 = a.max_steps;
    if (K64 < 1) K64 = 1;
    const uint32_t K = (uint32_t)K64;
    const bool valid = k < n_alive;
    uint32_t r = 0, count = 0;
    bool exited = false;
    V3 o = v3(0, 0, 0), d = v3(0, 0, 1);
    float t = 0.0f;
    if (valid) {
        r = alive[k];
        const float4 ro = r_o[r], rd = r_d[r];
        o = v3(ro.x, ro.y, ro.z);
        d = v3(rd.x, rd.y, rd.z);
        const V3 idir = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        t = ro.w;
        const uint32_t done = r_steps[r];
        const uint32_t budget = a.max_steps > done ? a.max_steps - done : 0u;
        const uint32_t kmax = K < budget ? K : budget;
        // one loop iteration per sample or per empty-voxel skip (the t
        // sequence of skip_empty + t += dt, without a nested loop whose trip
        // count differs between lanes)
        while (count < kmax) {
            const V3 pos = o + d * t;
            if (!box_contains(a, pos)) {
                exited = true;
                break;
            }
            if (occupied(a.grid, a.G, pos)) {
                t_buf[(size_t)count * n_alive + k] = t;
                ++count;
                t += a.dt;
            } else {
                t = advance_to_next_voxel(t, a.dt, pos, d, idir, a.G);
            }
        }
    }
    // the wave's samples are one contiguous range, written step-major
    const uint32_t wave_base = __shfl(wave_reserve(&counters[kCntSamples], count), 0);
    const uint32_t maxc = wave_max(count);
    StepMajor sm;
    for (uint32_t j = 0; j < maxc; ++j) {
        bool mine;
        const uint32_t slot = sm.next(j, count, mine);
        if (mine) {
            const V3 pos = o + d * t_buf[(size_t)j * n_alive + k];
            s_pos[wave_base + slot] = make_float4(pos.x, pos.y, pos.z, __uint_as_float(r));
        }
    }
    if (!valid) return;
    r_off[k] = wave_base;
    r_cnt[k] = count;
    r_o[r].w = t;
    r_flags[r] = exited ? 2u : 0u;
}

__global__ __launch_bounds__(kRayThreads) void composite_kernel(RenderArgs a, const uint32_t* __restrict__ alive,
                                                                const uint32_t* __restrict__ n_alive_dev,
                                                                const uint32_t* __restrict__ r_off,
                                                                const uint32_t* __restrict__ r_cnt,
                                                                const float4* __restrict__ s_pos,
                                                                const float4* __restrict__ s_out,
                                                                float4* __restrict__ r_d, float4* __restrict__ r_rgba,
                                                                float* __restrict__ r_depth,
                                                                uint32_t* __restrict__ r_steps,
                                                                const uint32_t* __restrict__ r_flags,
                                                                uint32_t* __restrict__ next_alive,
                                                                uint32_t* __restrict__ counters, uint32_t next_slot) {
    const uint32_t n_alive = *n_alive_dev;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0 && n_alive) {  // round statistics (march of this round is complete)
        counters[kCntEmitted] += counters[kCntSamples];
        counters[kCntRounds] += 1;
    }
    if ((k & ~63u) >= n_alive) return;
    const bool valid = k < n_alive;
    bool keep = false;
    uint32_t r = 0, used = 0;
    // the march's step-major layout of this wave's samples
    const uint32_t cnt = valid ? r_cnt[k] : 0u;
    const uint32_t off = r_off[k & ~63u];  // the wave's base (lane 0 is live)
    const uint32_t maxc = wave_max(cnt);
    if (valid) {
        r = alive[k];
        float4 acc = r_rgba[r];
        float4 rd = r_d[r];
        float depth = r_depth[r];
        uint32_t n = r_steps[r];
        const V3 cam_pos = v3(a.M[3], a.M[7], a.M[11]);
        const V3 cam_fwd = v3(a.M[2], a.M[6], a.M[10]);
        bool terminated = false;
        StepMajor sm;
        for (uint32_t j = 0; j < maxc; ++j) {
            bool mine;
            const uint32_t slot = sm.next(j, cnt, mine);
            if (!mine || terminated) continue;
            const float4 out = s_out[off + slot];
            const float4 sp = s_pos[off + slot];
            // composite_kernel_nerf (testbed_nerf.cu:853-999), render mode Shade
            const float T = 1.0f - acc.w;
            const float sigma = expf(out.w + a.density_bias);
            const float alpha = 1.0f - expf(-sigma * a.dt);
            const float weight = alpha * T;
            acc.x = acc.x + (1.0f / (1.0f + expf(-out.x))) * weight;
            acc.y = acc.y + (1.0f / (1.0f + expf(-out.y))) * weight;
            acc.z = acc.z + (1.0f / (1.0f + expf(-out.z))) * weight;
            acc.w = acc.w + weight;
            if (weight > rd.w) {
                rd.w = weight;
                depth = dot(cam_fwd, v3(sp.x, sp.y, sp.z) - cam_pos);
            }
            ++n;
            ++used;
            if (acc.w > 1.0f - a.min_transmittance) {
                const float s = acc.w;
                acc.x /= s; acc.y /= s; acc.z /= s; acc.w /= s;
                terminated = true;
            }
        }
        r_rgba[r] = acc;
        r_d[r] = rd;
        r_depth[r] = depth;
        r_steps[r] = n;
        keep = !terminated && !(r_flags[r] & 2u) && n < a.max_steps;
    }
    const uint32_t slot = wave_append(&counters[next_slot], keep);
    if (keep) next_alive[slot] = r;
    // statistics: composited samples this round
    uint32_t total = used;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) total += __shfl_xor(total, o);
    if ((threadIdx.x & 63u) == 0 && total) atomicAdd(&counters[kCntComposited], total);
}

__global__ void finalize_kernel(uint32_t n, const float4* __restrict__ r_rgba, const float* __restrict__ r_depth,
                                const uint32_t* __restrict__ r_steps, float4* __restrict__ rgba,
                                float* __restrict__ depth, uint32_t* __restrict__ steps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 acc = r_rgba[i];
    rgba[i] = acc.w > 0.001f ? acc : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (depth) depth[i] = acc.w > 0.2f ? r_depth[i] : 1e10f;
    if (steps) steps[i] = r_steps[i];
}

// ---- render buffer: shade, accumulate, tonemap ----------------------------
// (testbed_nerf.cu:1794-1827; render_buffer.cu:236-266, 264-350, 493-522),
// one thread per pixel, fused: the frame buffer of one sample index exists
// only in registers.
struct FrameArgs {
    float sample_count;     // spp: the accumulation's running-mean weight
    float exposure_scale;   // 2^exposure
    float bg[4];            // background, rgb already linear, weight
    float k[6];             // rational tonemap coefficients (ACES, Hable)
    uint32_t curve, output_srgb, linear_colors;
};

INSTA_DEV float srgb_to_linear(float c) { return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f); }
INSTA_DEV float linear_to_srgb(float c) { return c < 0.0031308f ? 12.92f * c : 1.055f * powf(c, 0.41666f) - 0.055f; }

INSTA_DEV V3 tonemap_curve(V3 x, const FrameArgs& a) {
    if (a.curve == INSTA_TONEMAP_IDENTITY) return x;
    x = v3(fmaxf(x.x, 0.0f), fmaxf(x.y, 0.0f), fmaxf(x.z, 0.0f));
    if (a.curve == INSTA_TONEMAP_REINHARD) {
        const float Y = 0.2126f * x.x + 0.7152f * x.y + 0.0722f * x.z;
        const float s = 1.0f / (Y + 1.0f);
        return v3(x.x * s, x.y * s, x.z * s);
    }
    float c[3] = {x.x, x.y, x.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float sq = c[i] * c[i];
        const float nom = sq * a.k[0] + a.k[1] * c[i] + a.k[2];
        const float den = a.k[3] * sq + a.k[4] * c[i] + a.k[5];
        c[i] = nom / den;
    }
    return v3(c[0], c[1], c[2]);
}

// rgba/out and depth/depth_out may be the same buffers (shaded in place)
__global__ void frame_buffer_kernel(uint32_t n, FrameArgs a, const float4* rgba, const float* depth,
                                    float4* __restrict__ accum, float4* out, float* depth_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // shade_kernel_nerf into a cleared frame buffer: tmp + 0 * (1 - tmp.w)
    float4 f = rgba[i];
    if (!a.linear_colors) f = make_float4(srgb_to_linear(f.x), srgb_to_linear(f.y), srgb_to_linear(f.z), f.w);
    if (depth_out) depth_out[i] = f.w > 0.2f && depth ? depth[i] : 0.0f;
    // accumulate_kernel, linear colour space (spp 0 starts from a cleared buffer)
    const float s = a.sample_count;
    float4 t = s == 0.0f ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : accum[i];
    t.x = (t.x * s + f.x) / (s + 1.0f);
    t.y = (t.y * s + f.y) / (s + 1.0f);
    t.z = (t.z * s + f.z) / (s + 1.0f);
    t.w = (t.w * s + f.w) / (s + 1.0f);
    // tonemap_kernel: background behind, exposure, curve, output colour space
    const float w = (1.0f - t.w) * a.bg[3];
    V3 c = v3(t.x + a.bg[0] * w, t.y + a.bg[1] * w, t.z + a.bg[2] * w);
    t.w = t.w + w;
    c = v3(c.x * a.exposure_scale, c.y * a.exposure_scale, c.z * a.exposure_scale);
    c = tonemap_curve(c, a);
    if (a.output_srgb) c = v3(linear_to_srgb(c.x), linear_to_srgb(c.y), linear_to_srgb(c.z));
    const float4 o = make_float4(c.x, c.y, c.z, t.w);
    accum[i] = o;  // the reference's tonemap_kernel writes its result back
    out[i] = o;
}

__global__ void clear_counters_kernel(uint32_t* counters, uint32_t first, uint32_t n) {
    const uint32_t i = threadIdx.x;
    if (i < n) counters[first + i] = 0;
}

// Before a round: clear the next live list's count and the round's sample
// count; the first round also keeps the initial live count.
__global__ void round_begin_kernel(uint32_t* counters, uint32_t next_slot, uint32_t first) {
    if (threadIdx.x != 0) return;
    if (first) counters[kCntAlive0] = counters[kCntAliveA];
    counters[next_slot] = 0;
    counters[kCntSamples] = 0;
}

// ---- host side of the Sobol pixel offset (random_val.cuh:272-322) --------
uint32_t h_bitrev(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}
uint32_t h_lk(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
uint32_t h_owen(uint32_t x, uint32_t seed) { return h_bitrev(h_lk(h_bitrev(x), seed)); }
uint32_t h_sobol(uint32_t index, uint32_t dim) {
    if (dim == 0) return h_bitrev(index);
    uint32_t v = 0x80000000u, x = 0;
    for (uint32_t bit = 0; bit < 32; ++bit) {
        if ((index >> bit) & 1u) x ^= v;
        v ^= v >> 1;
    }
    return x;
}
void h_ld_2d(uint32_t index, uint32_t seed, float out[2]) {
    const float S = (float)(1.0 / 4294967296.0);
    index = h_owen(index, seed);
    for (uint32_t i = 0; i < 2; ++i) out[i] = (float)h_owen(h_sobol(index, i), seed ^ (i + (seed << 6) + (seed >> 2))) * S;
}
void pixel_offset(uint32_t spp, float out[2]) {
    float a[2], b[2];
    h_ld_2d(0, 0xdeadbeefu, a);
    h_ld_2d(spp, 0xdeadbeefu, b);
    for (int i = 0; i < 2; ++i) {
        const float o = 0.5f - a[i] + b[i];
        out[i] = o - floorf(o);
    }
}

void prof_begin(insta_ctx* ctx, bool on, int cls, hipStream_t stream) {
    if (on) (void)hipEventRecord(ctx->prof_ev[2 * cls], stream);
}
void prof_end(insta_ctx* ctx, bool on, int cls, uint64_t items, hipStream_t stream) {
    if (!on) return;
    (void)hipEventRecord(ctx->prof_ev[2 * cls + 1], stream);
    ctx->prof_pending[cls] = true;
    ctx->prof_pending_items[cls] = items;
}
// after a stream synchronisation: fold the recorded intervals in
void prof_collect(insta_ctx* ctx, bool on) {
    if (!on) return;
    for (int c = 0; c < INSTA_KERNEL_CLASSES; ++c) {
        if (!ctx->prof_pending[c]) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ctx->prof_ev[2 * c], ctx->prof_ev[2 * c + 1]) == hipSuccess) {
            ctx->prof.ms[c] += ms;
            ctx->prof.launches[c] += 1;
            ctx->prof.items[c] += ctx->prof_pending_items[c];
        }
        ctx->prof_pending[c] = false;
    }
}

void free_render_ws(RenderWS& ws) {
    if (ws.external) {  // carved from memory owned elsewhere: forget the pointers only
        ws.r_o = ws.r_d = ws.r_rgba = nullptr;
        ws.r_depth = nullptr;
        ws.r_steps = ws.r_flags = ws.alive[0] = ws.alive[1] = ws.r_off = ws.r_cnt = nullptr;
        ws.s_pos = ws.s_cpos = ws.s_cdir = ws.s_out = nullptr;
        ws.s_hit = nullptr;
        ws.pixels = 0;
        ws.samples = 0;
        ws.bytes = 0;
        ws.external = false;
        return;
    }
    dev_free(ws.r_o); dev_free(ws.r_d); dev_free(ws.r_rgba); dev_free(ws.r_depth);
    dev_free(ws.r_steps); dev_free(ws.r_flags); dev_free(ws.alive[0]); dev_free(ws.alive[1]);
    dev_free(ws.r_off); dev_free(ws.r_cnt);
    dev_free(ws.s_pos); dev_free(ws.s_cpos); dev_free(ws.s_cdir); dev_free(ws.s_out);
    dev_free(ws.s_hit);
    ws.pixels = 0;
    ws.samples = 0;
    ws.bytes = 0;
}

uint64_t ws_sample_cap(uint32_t n_pixels, uint32_t spp) {
    // samples_per_pixel samples per pixel, between 1M and 64M samples (4 GiB)
    uint64_t cap = (uint64_t)n_pixels * spp;
    if (cap < (1u << 20)) cap = 1u << 20;
    if (cap > (64ull << 20)) cap = 64ull << 20;
    return cap;
}

// a render's workspace must already hold n_pixels (insta_reserve, or a
// binding onto caller memory): nothing is allocated per render
insta_status ensure_workspace(RenderWS& ws, uint32_t n_pixels) {
    if (ws.d_counters && n_pixels <= ws.pixels) return INSTA_OK;
    set_error("render: the workspace holds " + std::to_string(ws.pixels) + " pixels, the frame needs " +
              std::to_string(n_pixels) + ": reserve them with insta_reserve (render_pixels)");
    return INSTA_ERR_STATE;
}

}  // namespace

insta_status reserve_render_ws(RenderWS& ws, uint32_t n_pixels) {
    if (!ws.d_counters) {
        if (hipMalloc(reinterpret_cast<void**>(&ws.d_counters), kNumCounters * sizeof(uint32_t)) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&ws.h_counters), kNumCounters * sizeof(uint32_t),
                          hipHostMallocDefault) != hipSuccess) {
            set_error("render counters allocation failed");
            return INSTA_ERR_OOM;
        }
    }
    if (n_pixels <= ws.pixels) return INSTA_OK;
    if (ws.external) {
        set_error("render: bound workspace smaller than the request");
        return INSTA_ERR_INVALID;
    }
    free_render_ws(ws);
    const size_t n = n_pixels;
    const uint64_t cap = ws_sample_cap(n_pixels, ws.samples_per_pixel);
#define INSTA_ALLOC(ptr, bytes)                                           \
    if (hipMalloc(reinterpret_cast<void**>(&(ptr)), (bytes)) != hipSuccess) { \
        free_render_ws(ws);                                               \
        set_error("render workspace allocation failed");                 \
        return INSTA_ERR_OOM;                                             \
    }
    INSTA_ALLOC(ws.r_o, n * sizeof(float4));
    INSTA_ALLOC(ws.r_d, n * sizeof(float4));
    INSTA_ALLOC(ws.r_rgba, n * sizeof(float4));
    INSTA_ALLOC(ws.r_depth, n * sizeof(float));
    INSTA_ALLOC(ws.r_steps, n * sizeof(uint32_t));
    INSTA_ALLOC(ws.r_flags, n * sizeof(uint32_t));
    INSTA_ALLOC(ws.alive[0], n * sizeof(uint32_t));
    INSTA_ALLOC(ws.alive[1], n * sizeof(uint32_t));
    INSTA_ALLOC(ws.r_off, n * sizeof(uint32_t));
    INSTA_ALLOC(ws.r_cnt, n * sizeof(uint32_t));
    INSTA_ALLOC(ws.s_pos, cap * sizeof(float4));
    INSTA_ALLOC(ws.s_cpos, cap * sizeof(float4));
    INSTA_ALLOC(ws.s_cdir, cap * sizeof(float4));
    INSTA_ALLOC(ws.s_out, cap * sizeof(float4));
    INSTA_ALLOC(ws.s_hit, cap * sizeof(int32_t));
#undef INSTA_ALLOC
    ws.pixels = n_pixels;
    ws.samples = cap;
    ws.bytes = render_ws_bytes(n_pixels, ws.samples_per_pixel);
    return INSTA_OK;
}

constexpr uint64_t kWsAlign = 256;
constexpr uint64_t kWsBuffers = 15;

static_assert(kNumCounters == kRenderCounterWords, "render counter words");

uint64_t render_ws_bytes_cap(uint32_t n_pixels, uint64_t sample_cap) {
    return (uint64_t)n_pixels * (3 * sizeof(float4) + 7 * sizeof(uint32_t)) +
           sample_cap * (4 * sizeof(float4) + sizeof(int32_t)) + kWsBuffers * kWsAlign;
}

uint64_t render_ws_bytes(uint32_t n_pixels, uint32_t samples_per_pixel) {
    return render_ws_bytes_cap(n_pixels, ws_sample_cap(n_pixels, samples_per_pixel));
}

void bind_render_ws(RenderWS& ws, char* base, uint32_t n_pixels, uint32_t samples_per_pixel) {
    bind_render_ws_cap(ws, base, n_pixels, ws_sample_cap(n_pixels, samples_per_pixel));
    ws.samples_per_pixel = samples_per_pixel;
}

void bind_render_ws_cap(RenderWS& ws, char* base, uint32_t n_pixels, uint64_t sample_cap) {
    const uint64_t cap = sample_cap;
    const size_t n = n_pixels;
    uint64_t off = 0;
    auto carve = [&](auto*& ptr, uint64_t bytes) {
        ptr = reinterpret_cast<std::remove_reference_t<decltype(ptr)>>(base + off);
        off += (bytes + kWsAlign - 1) / kWsAlign * kWsAlign;
    };
    carve(ws.r_o, n * sizeof(float4));
    carve(ws.r_d, n * sizeof(float4));
    carve(ws.r_rgba, n * sizeof(float4));
    carve(ws.r_depth, n * sizeof(float));
    carve(ws.r_steps, n * sizeof(uint32_t));
    carve(ws.r_flags, n * sizeof(uint32_t));
    carve(ws.alive[0], n * sizeof(uint32_t));
    carve(ws.alive[1], n * sizeof(uint32_t));
    carve(ws.r_off, n * sizeof(uint32_t));
    carve(ws.r_cnt, n * sizeof(uint32_t));
    carve(ws.s_pos, cap * sizeof(float4));
    carve(ws.s_cpos, cap * sizeof(float4));
    carve(ws.s_cdir, cap * sizeof(float4));
    carve(ws.s_out, cap * sizeof(float4));
    carve(ws.s_hit, cap * sizeof(int32_t));
    ws.pixels = n_pixels;
    ws.samples = cap;
    ws.bytes = off;
    ws.external = true;
}

namespace {

RenderArgs render_args(const insta_ctx* ctx, const insta_camera* cam, uint32_t width, uint32_t height,
                       int32_t occupancy, uint32_t spp) {
    RenderArgs a;
    for (int i = 0; i < 12; ++i) a.M[i] = cam->matrix[i];
    a.cx = cam->screen_center[0];
    a.cy = cam->screen_center[1];
    a.fx = cam->focal_px[0];
    a.fy = cam->focal_px[1];
    float off[2];
    pixel_offset(cam->snap_to_pixel_centers ? 0u : spp, off);
    a.off_x = off[0];
    a.off_y = off[1];
    a.width = width;
    a.height = height;
    a.spp = spp;
    a.G = ctx->cfg.grid_size;
    a.max_steps = ctx->cfg.max_steps;
    a.dt = ctx->cfg.dt;
    a.min_transmittance = ctx->cfg.min_transmittance;
    a.density_bias = ctx->cfg.density_logit_bias;
    for (int i = 0; i < 3; ++i) {
        a.lo[i] = ctx->render_lo[i];
        a.hi[i] = ctx->render_hi[i];
    }
    a.grid = ctx->grid_lin(occupancy < 0 ? 0u : (uint32_t)occupancy + 1u);
    return a;
}

// One march/deform/network/composite round over live list `cur`, with the
// live count read on the device.  `max_alive` bounds the live count (it
// sizes the ray grids).
insta_status enqueue_round(insta_ctx* ctx, RenderWS& ws, const RenderArgs& a, int32_t frame, uint32_t cur,
                           bool first, uint32_t max_alive, bool prof, hipStream_t stream) {
    uint32_t* cnt = ws.d_counters;
    const uint32_t next = 1u - cur;
    const uint32_t cur_slot = cur == 0 ? kCntAliveA : kCntAliveB;
    const uint32_t next_slot = next == 0 ? kCntAliveA : kCntAliveB;
    hipLaunchKernelGGL(round_begin_kernel, dim3(1), dim3(64), 0, stream, cnt, next_slot, first ? 1u : 0u);
    const dim3 rgrid((max_alive + kRayThreads - 1) / kRayThreads);
    prof_begin(ctx, prof, INSTA_KERNEL_MARCH, stream);
    hipLaunchKernelGGL(march_kernel, rgrid, dim3(kRayThreads), 0, stream, a, ws.alive[cur], cnt + cur_slot,
                       (uint64_t)ws.samples, ws.r_o, ws.r_d, ws.r_steps, ws.r_flags, ws.r_off, ws.r_cnt,
                       reinterpret_cast<float*>(ws.s_hit), ws.s_pos, cnt);
    INSTA_HIP(hipGetLastError());
    prof_end(ctx, prof, INSTA_KERNEL_MARCH, 0, stream);
    // at most max_alive * K <= sample capacity samples this round
    const uint32_t n_max = (uint32_t)ws.samples;
    prof_begin(ctx, prof, INSTA_KERNEL_DEFORM, stream);
    insta_status st = launch_deform_soa(ctx, ws, frame, cnt + kCntSamples, n_max, cnt + kCntMisses, stream);
    if (st != INSTA_OK) return st;
    prof_end(ctx, prof, INSTA_KERNEL_DEFORM, 0, stream);
    prof_begin(ctx, prof, INSTA_KERNEL_NETWORK, stream);
    st = launch_network_samples(ctx, cnt + kCntSamples, n_max, ws.s_cpos, ws.s_cdir, ws.s_out, stream);
    if (st != INSTA_OK) return st;
    prof_end(ctx, prof, INSTA_KERNEL_NETWORK, 0, stream);
    prof_begin(ctx, prof, INSTA_KERNEL_COMPOSITE, stream);
    hipLaunchKernelGGL(composite_kernel, rgrid, dim3(kRayThreads), 0, stream, a, ws.alive[cur], cnt + cur_slot,
                       ws.r_off, ws.r_cnt, ws.s_pos, ws.s_out, ws.r_d, ws.r_rgba, ws.r_depth, ws.r_steps,
                       ws.r_flags, ws.alive[next], cnt, next_slot);
    INSTA_HIP(hipGetLastError());
    prof_end(ctx, prof, INSTA_KERNEL_COMPOSITE, 0, stream);
    return INSTA_OK;
}

insta_status enqueue_finalize(insta_ctx* ctx, RenderWS& ws, uint32_t n_pix, float* rgba, float* depth,
                              uint32_t* steps, bool prof, hipStream_t stream) {
    prof_begin(ctx, prof, INSTA_KERNEL_FINALIZE, stream);
    hipLaunchKernelGGL(finalize_kernel, dim3((n_pix + kRayThreads - 1) / kRayThreads), dim3(kRayThreads), 0, stream,
                       n_pix, ws.r_rgba, ws.r_depth, ws.r_steps, reinterpret_cast<float4*>(rgba), depth, steps);
    INSTA_HIP(hipGetLastError());
    prof_end(ctx, prof, INSTA_KERNEL_FINALIZE, n_pix, stream);
    return INSTA_OK;
}

// profiled item counts, known once the counters are on the host
void prof_items(insta_ctx* ctx, bool prof, uint32_t alive, uint32_t samples) {
    if (!prof) return;
    ctx->prof_pending_items[INSTA_KERNEL_MARCH] = alive;
    ctx->prof_pending_items[INSTA_KERNEL_COMPOSITE] = alive;
    ctx->prof_pending_items[INSTA_KERNEL_DEFORM] = samples;
    ctx->prof_pending_items[INSTA_KERNEL_NETWORK] = samples;
}

}  // namespace

insta_status enqueue_render(insta_ctx* ctx, RenderWS& ws, const insta_camera* cam, uint32_t width, uint32_t height,
                            int32_t frame, int32_t occupancy, uint32_t spp, uint32_t rounds, float* rgba,
                            float* depth, uint32_t* steps, hipStream_t stream) {
    const uint32_t n_pix = width * height;
    const bool prof = ctx->profiling && &ws == &ctx->ws && rounds == 1;
    insta_status st = ensure_workspace(ws, n_pix);
    if (st != INSTA_OK) return st;
    const RenderArgs a = render_args(ctx, cam, width, height, occupancy, spp);
    uint32_t* cnt = ws.d_counters;
    hipLaunchKernelGGL(clear_counters_kernel, dim3(1), dim3(64), 0, stream, cnt, 0u, (uint32_t)kNumCounters);
    prof_begin(ctx, prof, INSTA_KERNEL_INIT, stream);
    hipLaunchKernelGGL(init_rays_kernel, dim3((tiled_threads(width, height) + kRayThreads - 1) / kRayThreads),
                       dim3(kRayThreads), 0, stream, a,
                       ws.r_o, ws.r_d, ws.r_rgba, ws.r_depth, ws.r_steps, ws.alive[0], cnt);
    INSTA_HIP(hipGetLastError());
    prof_end(ctx, prof, INSTA_KERNEL_INIT, n_pix, stream);
    for (uint32_t r = 0; r < rounds; ++r) {
        st = enqueue_round(ctx, ws, a, frame, r & 1u, r == 0, n_pix, prof, stream);
        if (st != INSTA_OK) return st;
    }
    return enqueue_finalize(ctx, ws, n_pix, rgba, depth, steps, prof, stream);
}

insta_status finish_render(insta_ctx* ctx, RenderWS& ws, const insta_camera* cam, uint32_t width, uint32_t height,
                           int32_t frame, int32_t occupancy, uint32_t spp, uint32_t rounds, float* rgba, float* depth,
                           uint32_t* steps, hipStream_t stream) {
    const uint32_t n_pix = width * height;
    const bool prof = ctx->profiling && &ws == &ctx->ws && rounds == 1;
    uint32_t* cnt = ws.d_counters;
    INSTA_HIP(hipMemcpyAsync(ws.h_counters, cnt, kNumCounters * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    INSTA_HIP(hipStreamSynchronize(stream));
    prof_items(ctx, prof, ws.h_counters[kCntAlive0], ws.h_counters[kCntSamples]);
    prof_collect(ctx, prof);
    uint32_t cur = rounds & 1u;
    uint32_t n_alive = ws.h_counters[cur == 0 ? kCntAliveA : kCntAliveB];
    if (n_alive) {
        // rays left after the enqueued rounds: one round per host check
        const RenderArgs a = render_args(ctx, cam, width, height, occupancy, spp);
        while (n_alive) {
            insta_status st = enqueue_round(ctx, ws, a, frame, cur, false, n_alive, prof, stream);
            if (st != INSTA_OK) return st;
            INSTA_HIP(hipMemcpyAsync(ws.h_counters, cnt, kNumCounters * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                     stream));
            INSTA_HIP(hipStreamSynchronize(stream));
            prof_items(ctx, prof, n_alive, ws.h_counters[kCntSamples]);
            prof_collect(ctx, prof);
            cur = 1u - cur;
            n_alive = ws.h_counters[cur == 0 ? kCntAliveA : kCntAliveB];
        }
        // the image written by the enqueued finalize was partial
        insta_status st = enqueue_finalize(ctx, ws, n_pix, rgba, depth, steps, false, stream);
        if (st != INSTA_OK) return st;
    }
    ws.stats = insta_render_stats{};
    ws.stats.n_rays = n_pix;
    ws.stats.n_rays_alive0 = ws.h_counters[kCntAlive0];
    ws.stats.n_samples = ws.h_counters[kCntComposited];
    ws.stats.n_samples_emitted = ws.h_counters[kCntEmitted];
    ws.stats.n_iterations = ws.h_counters[kCntRounds];
    ws.stats.n_deform_misses = ws.h_counters[kCntMisses];
    return INSTA_OK;
}

insta_status render(insta_ctx* ctx, RenderWS& ws, const insta_camera* cam, uint32_t width, uint32_t height,
                    int32_t frame, int32_t occupancy, uint32_t spp, float* rgba, float* depth, uint32_t* steps,
                    hipStream_t stream) {
    insta_status st = enqueue_render(ctx, ws, cam, width, height, frame, occupancy, spp, 1, rgba, depth, steps, stream);
    if (st != INSTA_OK) return st;
    return finish_render(ctx, ws, cam, width, height, frame, occupancy, spp, 1, rgba, depth, steps, stream);
}

insta_status set_profiling(insta_ctx* ctx, int level) {
    const bool on = level != 0;
    if (on && !ctx->prof_ev[0]) {
        for (int i = 0; i < 2 * INSTA_KERNEL_CLASSES; ++i) INSTA_HIP(hipEventCreate(&ctx->prof_ev[i]));
    }
    ctx->profiling = on;
    if (on) {
        ctx->prof = insta_profile{};
        for (int c = 0; c < INSTA_KERNEL_CLASSES; ++c) ctx->prof_pending[c] = false;
    }
    if (level == 2 && !ctx->d_trav_counts) {
        INSTA_HIP(hipMalloc(reinterpret_cast<void**>(&ctx->d_trav_counts), kTravCounters * sizeof(unsigned long long)));
        INSTA_HIP(hipMemset(ctx->d_trav_counts, 0, kTravCounters * sizeof(unsigned long long)));
    } else if (level != 2 && ctx->d_trav_counts) {
        INSTA_HIP(hipDeviceSynchronize());  // queued counting kernels may still write
        dev_free(ctx->d_trav_counts);
    }
    return INSTA_OK;
}

// host half of the render buffer: the per-call constants of
// frame_buffer_kernel (background to linear, exposure scale and the
// tonemap curve's rational coefficients, render_buffer.cu:264-325, 493-506)
namespace {
float h_srgb_to_linear(float c) { return c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f); }
}

insta_status launch_frame_buffer(const float4* rgba, const float* depth, uint32_t n, uint32_t spp,
                                 const insta_frame_config& fc, float4* accum, float4* out, float* depth_out,
                                 hipStream_t stream) {
    if (n == 0) return INSTA_OK;
    FrameArgs a{};
    a.sample_count = (float)spp;
    a.exposure_scale = std::pow(2.0f, fc.exposure);
    for (int c = 0; c < 3; ++c) a.bg[c] = h_srgb_to_linear(fc.background[c]);
    a.bg[3] = fc.background[3];
    a.curve = fc.tonemap_curve;
    a.output_srgb = fc.output_srgb;
    a.linear_colors = fc.linear_colors;
    if (fc.tonemap_curve == INSTA_TONEMAP_ACES) {
        // Narkowicz's ACES fit with the 0.6 pre-exposure folded in
        a.k[0] = 0.6f * 0.6f * 2.51f; a.k[1] = 0.6f * 0.03f; a.k[2] = 0.0f;
        a.k[3] = 0.6f * 0.6f * 2.43f; a.k[4] = 0.6f * 0.59f; a.k[5] = 0.14f;
    } else if (fc.tonemap_curve == INSTA_TONEMAP_HABLE) {
        // Hable's filmic curve, white point 11.2, exposure bias 2 folded in
        const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f, W = 11.2f;
        float k0 = A * F - A * E, k1 = C * B * F - B * E, k2 = 0.0f, k3 = A * F, k4 = B * F, k5 = D * F * F;
        const float nom = k0 * (W * W) + k1 * W + k2;
        const float den = k3 * (W * W) + k4 * W + k5;
        const float white = den / nom;
        a.k[0] = 4.0f * k0 * white; a.k[1] = 2.0f * k1 * white; a.k[2] = k2 * white;
        a.k[3] = 4.0f * k3; a.k[4] = 2.0f * k4; a.k[5] = k5;
    }
    hipLaunchKernelGGL(frame_buffer_kernel, dim3((n + kRayThreads - 1) / kRayThreads), dim3(kRayThreads), 0, stream,
                       n, a, rgba, depth, accum, out, depth_out);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

void free_workspace(insta_ctx* ctx) {
    dev_free(ctx->d_trav_counts);
    if (ctx->aux_stream) (void)hipStreamDestroy(ctx->aux_stream);
    ctx->aux_stream = nullptr;
    for (int i = 0; i < 2 * INSTA_KERNEL_CLASSES; ++i)
        if (ctx->prof_ev[i]) {
            (void)hipEventDestroy(ctx->prof_ev[i]);
            ctx->prof_ev[i] = nullptr;
        }
    release_render_ws(ctx->ws);
}

void release_render_ws(RenderWS& ws) {
    free_render_ws(ws);
    ws.external = false;
    dev_free(ws.d_counters);
    if (ws.h_counters) (void)hipHostFree(ws.h_counters);
    ws.h_counters = nullptr;
}

}  // namespace insta
