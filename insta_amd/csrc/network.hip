// network.hip -- hash-grid encoding (forward, backward scatter-add) and the
// fused density/colour MLP evaluation, FP32 path.
//
// Composition follows NerfNetwork::inference_mixed_precision_impl
// (instant-ngp/include/neural-graphics-primitives/nerf_network.h:159-232):
//   features = HashGrid(pos) [| extra dims]
//   d        = density MLP(features)            (16 outputs, bias-free, ReLU hidden)
//   rgb      = colour MLP([d | SH_deg4(dir)])   (3 outputs)
//   output   = (rgb logits, d[0] = density logit)
// The hash-grid convention is stated in encoding.hpp.
#include "context.hpp"
#include "encoding.hpp"

namespace insta {
namespace {

constexpr uint32_t kNetThreads = 128;
constexpr uint32_t kEncThreads = 256;

template <int F>
__global__ __launch_bounds__(kEncThreads) void encode_kernel(const float* __restrict__ table, GridParams gp,
                                                             uint32_t n_levels, const float* __restrict__ pos,
                                                             uint32_t n, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 x = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    float* o = out + (size_t)i * n_levels * F;
    for (uint32_t l = 0; l < n_levels; ++l) {
        float acc[F];
        encode_level<F>(table, gp, l, x, acc);
#pragma unroll
        for (int f = 0; f < F; ++f) o[l * F + f] = acc[f];
    }
}

// Backward: each sample scatters dL/dfeature * weight into its 8 corners per
// level with hardware FP32 atomics (global_atomic_add_f32).
template <int F>
__global__ __launch_bounds__(kEncThreads) void encode_backward_kernel(GridParams gp, uint32_t n_levels,
                                                                      const float* __restrict__ pos, uint32_t n,
                                                                      const float* __restrict__ dL,
                                                                      float* __restrict__ grad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float xs[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    for (uint32_t l = 0; l < n_levels; ++l) {
        const float sc = gp.scale[l];
        float frac[3];
        uint32_t g[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float p = xs[d] * sc + 0.5f;
            const float fl = floorf(p);
            frac[d] = p - fl;
            g[d] = (uint32_t)(int32_t)fl;
        }
        float gl[F];
#pragma unroll
        for (int f = 0; f < F; ++f) gl[f] = dL[((size_t)i * n_levels + l) * F + f];
        const bool dense = (gp.dense_mask >> l) & 1u;
        const uint32_t res = gp.res[l], size = gp.size[l];
        float* base = grad + (size_t)gp.offset[l] * F;
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) {
            float w = 1.0f;
            uint32_t cg[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const bool up = (c >> d) & 1u;
                w = w * (up ? frac[d] : 1.0f - frac[d]);
                cg[d] = g[d] + (up ? 1u : 0u);
            }
            const uint32_t idx = grid_index(cg, dense, res, size);
#pragma unroll
            for (int f = 0; f < F; ++f) atomicAdd(base + (size_t)idx * F + f, w * gl[f]);
        }
    }
}

struct NetArgs {
    const float* params;   // MLP weights followed by the hash tables
    const float* table;
    const float* frame_extra;  // [max_frames][X] per-frame extra dims
    GridParams gp;
    uint32_t density_hidden, rgb_hidden;
    uint32_t n_frames;         // rows of frame_extra
};

constexpr int kWidth = 64;
constexpr int kDensityOut = 16;
constexpr int kRgbIn = kDensityOut + 16;

// extra: the sample's X extra dims (null: all 1.0)
// First density layer, level by level: each level's F features go straight
// into the 64 accumulators, so a wide input (16 levels x 8 features) never
// lives as one unrolled vector.  Every output still sums its inputs in index
// order with one FMA each, so the result equals dense<DIN, 64, true> bit for bit.
template <int L, int F, int X>
INSTA_DEV void first_layer_by_level(const NetArgs& a, V3 pos, const float* extra, float h[kWidth]) {
    constexpr int DIN = L * F + X;
    const float* W = a.params;
#pragma unroll
    for (int o = 0; o < kWidth; ++o) h[o] = 0.0f;
#pragma unroll 1
    for (int l = 0; l < L; ++l) {
        float acc[F];
        encode_level<F>(a.table, a.gp, (uint32_t)l, pos, acc);
#pragma unroll
        for (int o = 0; o < kWidth; ++o) {
#pragma unroll
            for (int f = 0; f < F; ++f) h[o] = __builtin_fmaf(W[o * DIN + l * F + f], acc[f], h[o]);
        }
    }
#pragma unroll
    for (int e = 0; e < X; ++e) {
        const float xe = extra ? extra[e] : 1.0f;
#pragma unroll
        for (int o = 0; o < kWidth; ++o) h[o] = __builtin_fmaf(W[o * DIN + L * F + e], xe, h[o]);
    }
#pragma unroll
    for (int o = 0; o < kWidth; ++o) h[o] = h[o] > 0.0f ? h[o] : 0.0f;
}

template <int L, int F, int X>
INSTA_DEV void network_one(const NetArgs& a, V3 pos, V3 dir_w, const float* extra, float out[4]) {
    constexpr int DIN = L * F + X;
    float h[kWidth], g[kWidth];
    const float* W = a.params;
    if constexpr (F >= 8) {
        first_layer_by_level<L, F, X>(a, pos, extra, h);
    } else {
        float x[DIN];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            float acc[F];
            encode_level<F>(a.table, a.gp, (uint32_t)l, pos, acc);
#pragma unroll
            for (int f = 0; f < F; ++f) x[l * F + f] = acc[f];
        }
#pragma unroll
        for (int e = 0; e < X; ++e) x[L * F + e] = extra ? extra[e] : 1.0f;
        dense<DIN, kWidth, true>(W, x, h);
    }
    W += kWidth * DIN;
    for (uint32_t k = 1; k < a.density_hidden; ++k) {
        dense<kWidth, kWidth, true>(W, h, g);
#pragma unroll
        for (int i = 0; i < kWidth; ++i) h[i] = g[i];
        W += kWidth * kWidth;
    }
    float r[kRgbIn];
    dense<kWidth, kDensityOut, false>(W, h, r);
    W += kDensityOut * kWidth;
    sh16(dir_w.x * 2.0f - 1.0f, dir_w.y * 2.0f - 1.0f, dir_w.z * 2.0f - 1.0f, r + kDensityOut);
    dense<kRgbIn, kWidth, true>(W, r, h);
    W += kWidth * kRgbIn;
    for (uint32_t k = 1; k < a.rgb_hidden; ++k) {
        dense<kWidth, kWidth, true>(W, h, g);
#pragma unroll
        for (int i = 0; i < kWidth; ++i) h[i] = g[i];
        W += kWidth * kWidth;
    }
    float rgb[3];
    dense<kWidth, 3, false>(W, h, rgb);
    out[0] = rgb[0];
    out[1] = rgb[1];
    out[2] = rgb[2];
    out[3] = r[0];
}

template <int L, int F, int X>
__global__ __launch_bounds__(kNetThreads) void network_aos_kernel(NetArgs a, const float* __restrict__ coords,
                                                                  uint32_t n, uint32_t stride,
                                                                  float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* c = coords + (size_t)i * stride;
    float o[4];
    network_one<L, F, X>(a, v3(c[0], c[1], c[2]), v3(c[8], c[9], c[10]), c + stride - X, o);
    reinterpret_cast<float4*>(out)[i] = make_float4(o[0], o[1], o[2], o[3]);
}

// the extra dims come from the selector in cdir.w (0: all 1.0, 1 + f: frame f)
template <int L, int F, int X>
__global__ __launch_bounds__(kNetThreads) void network_soa_kernel(NetArgs a, const uint32_t* __restrict__ n_dev,
                                                                  const float4* __restrict__ cpos,
                                                                  const float4* __restrict__ cdir,
                                                                  float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    const float4 p = cpos[i], d = cdir[i];
    const uint32_t sel = X ? __float_as_uint(d.w) : kExtraOnes;
    const float* extra =
        sel != kExtraOnes && sel <= a.n_frames ? a.frame_extra + (size_t)(sel - 1u) * X : nullptr;
    float o[4];
    network_one<L, F, X>(a, v3(p.x, p.y, p.z), v3(d.x, d.y, d.z), extra, o);
    out[i] = make_float4(o[0], o[1], o[2], o[3]);
}

NetArgs net_args(const insta_ctx* ctx) {
    NetArgs a;
    a.params = net_params(ctx);
    a.table = net_params(ctx) + ctx->layout.n_mlp_params;
    a.frame_extra = ctx->d_frame_extra;
    a.n_frames = ctx->cfg.max_frames;
    a.gp = grid_params(ctx);
    a.density_hidden = ctx->cfg.density_hidden_layers;
    a.rgb_hidden = ctx->cfg.rgb_hidden_layers;
    return a;
}

}  // namespace

GridParams grid_params(const insta_ctx* ctx) {
    GridParams gp{};
    for (uint32_t l = 0; l < ctx->cfg.n_levels; ++l) {
        gp.scale[l] = ctx->layout.level_scale[l];
        gp.res[l] = ctx->layout.level_resolution[l];
        gp.offset[l] = ctx->layout.level_offset[l];
        gp.size[l] = ctx->layout.level_size[l];
        if (ctx->layout.level_dense[l]) gp.dense_mask |= 1u << l;
    }
    return gp;
}

bool network_supported(const insta_config& c) {
    const bool shape = c.n_neurons == 64 && c.density_out == 16 && c.sh_degree == 4 &&
                       (c.n_extra_dims == 0 || c.n_extra_dims == 16) &&
                       c.density_hidden_layers >= 1 && c.rgb_hidden_layers >= 1;
    // F = 2 (the workload shapes) or the reference's configs/nerf/main.json
    // grid (16 levels of 8 features) on this per-sample FP32 path
    const bool grid = (c.n_features_per_level == 2 && (c.n_levels == 8 || c.n_levels == 16)) ||
                      (c.n_features_per_level == 8 && c.n_levels == 16);
    return shape && grid;
}

#define INSTA_DISPATCH_LF(L_, F_, X_, KERNEL, ...)                                                  \
    if (ctx->cfg.n_levels == L_ && ctx->cfg.n_features_per_level == F_ && ctx->cfg.n_extra_dims == X_) { \
        hipLaunchKernelGGL((KERNEL<L_, F_, X_>), __VA_ARGS__);                                       \
        INSTA_HIP(hipGetLastError());                                                                \
        return INSTA_OK;                                                                             \
    }

insta_status launch_network_aos(insta_ctx* ctx, const float* coords, uint32_t n, uint32_t stride, float* out,
                                hipStream_t stream) {
    if (n == 0) return INSTA_OK;
    const NetArgs a = net_args(ctx);
    const dim3 grid((n + kNetThreads - 1) / kNetThreads), block(kNetThreads);
    INSTA_DISPATCH_LF(8, 2, 0, network_aos_kernel, grid, block, 0, stream, a, coords, n, stride, out)
    INSTA_DISPATCH_LF(16, 2, 0, network_aos_kernel, grid, block, 0, stream, a, coords, n, stride, out)
    INSTA_DISPATCH_LF(8, 2, 16, network_aos_kernel, grid, block, 0, stream, a, coords, n, stride, out)
    INSTA_DISPATCH_LF(16, 2, 16, network_aos_kernel, grid, block, 0, stream, a, coords, n, stride, out)
    INSTA_DISPATCH_LF(16, 8, 0, network_aos_kernel, grid, block, 0, stream, a, coords, n, stride, out)
    INSTA_DISPATCH_LF(16, 8, 16, network_aos_kernel, grid, block, 0, stream, a, coords, n, stride, out)
    set_error("network: unsupported (n_levels, n_features_per_level)");
    return INSTA_ERR_UNSUPPORTED;
}

insta_status launch_network_soa(insta_ctx* ctx, const uint32_t* n_dev, uint32_t n_max, const float4* cpos,
                                const float4* cdir, float4* out, hipStream_t stream) {
    if (n_max == 0) return INSTA_OK;
    const NetArgs a = net_args(ctx);
    const dim3 grid((n_max + kNetThreads - 1) / kNetThreads), block(kNetThreads);
    INSTA_DISPATCH_LF(8, 2, 0, network_soa_kernel, grid, block, 0, stream, a, n_dev, cpos, cdir, out)
    INSTA_DISPATCH_LF(16, 2, 0, network_soa_kernel, grid, block, 0, stream, a, n_dev, cpos, cdir, out)
    INSTA_DISPATCH_LF(8, 2, 16, network_soa_kernel, grid, block, 0, stream, a, n_dev, cpos, cdir, out)
    INSTA_DISPATCH_LF(16, 2, 16, network_soa_kernel, grid, block, 0, stream, a, n_dev, cpos, cdir, out)
    INSTA_DISPATCH_LF(16, 8, 0, network_soa_kernel, grid, block, 0, stream, a, n_dev, cpos, cdir, out)
    INSTA_DISPATCH_LF(16, 8, 16, network_soa_kernel, grid, block, 0, stream, a, n_dev, cpos, cdir, out)
    set_error("network: unsupported (n_levels, n_features_per_level)");
    return INSTA_ERR_UNSUPPORTED;
}

insta_status launch_encode(insta_ctx* ctx, const float* pos, uint32_t n, float* out, hipStream_t stream) {
    if (n == 0) return INSTA_OK;
    const float* table = ctx->d_params + ctx->layout.n_mlp_params;
    const GridParams gp = grid_params(ctx);
    const dim3 grid((n + kEncThreads - 1) / kEncThreads), block(kEncThreads);
    switch (ctx->cfg.n_features_per_level) {
        case 1: hipLaunchKernelGGL(encode_kernel<1>, grid, block, 0, stream, table, gp, ctx->cfg.n_levels, pos, n, out); break;
        case 2: hipLaunchKernelGGL(encode_kernel<2>, grid, block, 0, stream, table, gp, ctx->cfg.n_levels, pos, n, out); break;
        case 4: hipLaunchKernelGGL(encode_kernel<4>, grid, block, 0, stream, table, gp, ctx->cfg.n_levels, pos, n, out); break;
        case 8: hipLaunchKernelGGL(encode_kernel<8>, grid, block, 0, stream, table, gp, ctx->cfg.n_levels, pos, n, out); break;
        default: set_error("encode: n_features_per_level must be 1, 2, 4 or 8"); return INSTA_ERR_UNSUPPORTED;
    }
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_encode_backward(insta_ctx* ctx, const float* pos, uint32_t n, const float* dL, float* grad,
                                    hipStream_t stream) {
    if (n == 0) return INSTA_OK;
    const GridParams gp = grid_params(ctx);
    const dim3 grid((n + kEncThreads - 1) / kEncThreads), block(kEncThreads);
    switch (ctx->cfg.n_features_per_level) {
        case 1: hipLaunchKernelGGL(encode_backward_kernel<1>, grid, block, 0, stream, gp, ctx->cfg.n_levels, pos, n, dL, grad); break;
        case 2: hipLaunchKernelGGL(encode_backward_kernel<2>, grid, block, 0, stream, gp, ctx->cfg.n_levels, pos, n, dL, grad); break;
        case 4: hipLaunchKernelGGL(encode_backward_kernel<4>, grid, block, 0, stream, gp, ctx->cfg.n_levels, pos, n, dL, grad); break;
        case 8: hipLaunchKernelGGL(encode_backward_kernel<8>, grid, block, 0, stream, gp, ctx->cfg.n_levels, pos, n, dL, grad); break;
        default: set_error("encode: n_features_per_level must be 1, 2, 4 or 8"); return INSTA_ERR_UNSUPPORTED;
    }
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

// canonical samples (cpos, cdir) -> out through the context's inference path
insta_status launch_network_samples(insta_ctx* ctx, const uint32_t* n_dev, uint32_t n_max, const float4* cpos,
                                    const float4* cdir, float4* out, hipStream_t stream) {
    return mfma_network_supported(ctx->cfg) ? launch_network_mfma_soa(ctx, n_dev, n_max, cpos, cdir, out, stream)
                                            : launch_network_soa(ctx, n_dev, n_max, cpos, cdir, out, stream);
}

}  // namespace insta
