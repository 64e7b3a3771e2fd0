// density.hip -- density-grid update of training: the occupancy grid as an
// EMA of the network's optical thickness (Testbed::update_density_grid_nerf,
// instant-ngp/src/testbed_nerf.cu:2774-2861, one cascade, one step):
//   samples   generate_grid_samples_nerf_nonuniform (487-516): n_uniform
//             cells at any density and n_nonuniform cells above 0.01 (up to
//             10 probes of a multiplicative-hash cell sequence), a random
//             position inside each cell; both sets read the same PCG32 state,
//             which then advances by 2^32 (pcg32's default jump)
//   warp      each position to canonical space through a random posed frame
//             (warp_random_samples, src/core.cu:528-579)
//   splat     density MLP at the canonical position (extra dims all 1.0, as
//             extract_expressions_random writes, nerf_network.h:107-127),
//             optical thickness
//             exp(logit + bias) * dt, atomicMax on the float bits into the
//             cell (splat_grid_samples_nerf_max_nearest_neighbor, 518-534)
//   ema       grid = max(grid * decay, new) unless negative
//             (ema_grid_samples_nerf, 554-577)
//   bitfield  cell occupied when grid > min(0.01, mean of max(grid, 0))
//             (update_density_grid_mean_and_bitfield + grid_to_bitfield,
//             2863-2878, 585-609), written into an occupancy slot
// The mean is a two-stage fixed-order reduction, so the bitfield is a
// deterministic function of the grid.
#include "context.hpp"
#include "encoding.hpp"

namespace insta {

struct DensityGridState {
    float* grid = nullptr;      // [G^3] Morton order
    float* tmp = nullptr;       // [G^3] this update's splats
    float* pos = nullptr;       // [cap][3]
    uint32_t* idx = nullptr;    // [cap]
    float* partial = nullptr;   // [G^3 / kReduce] block sums
    float* mean = nullptr;      // [1]
    float* h_mean = nullptr;    // pinned mirror
    uint32_t cap = 0;
    uint32_t step = 0;
    uint32_t marked_n = 0;      // training images at the last untrained-cell marking
    Pcg32 rng{0, 0};
};

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kReduce = 1024;    // cells per first-stage block
constexpr float kMinOpticalThickness = 0.01f;  // NERF_MIN_OPTICAL_THICKNESS (testbed_nerf.cu:63)

__global__ void dg_fill_kernel(float* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// mark_untrained_density_grid (testbed_nerf.cu:391-431): -1 for a cell that
// no training camera sees (centre widened by the half-diagonal against each
// camera's centred frustum), 0 for a seen cell that was untrained (or always
// with `clear`); one thread per cell, Morton order, one cascade
__global__ void dg_mark_kernel(float* __restrict__ grid, uint32_t G, const TrainImage* __restrict__ meta,
                               uint32_t n_images, float half_x, float half_y, uint32_t clear) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G * G * G) return;
    const float x = (float)compact3(i), y = (float)compact3(i >> 1), z = (float)compact3(i >> 2);
    const float px = ((x + 0.5f) / (float)G - 0.5f) * 1.0f + 0.5f;
    const float py = ((y + 0.5f) / (float)G - 0.5f) * 1.0f + 0.5f;
    const float pz = ((z + 0.5f) / (float)G - 0.5f) * 1.0f + 0.5f;
    const float radius = 0.5f * 1.73205080757f * 1.0f / (float)G;
    bool seen = false;
    for (uint32_t j = 0; j < n_images && !seen; ++j) {
        const TrainImage& c = meta[j];
        const float lx = px - c.M[3], ly = py - c.M[7], lz = pz - c.M[11];
        const float cx = lx * c.M[0] + ly * c.M[4] + lz * c.M[8];
        const float cy = lx * c.M[1] + ly * c.M[5] + lz * c.M[9];
        const float cz = lx * c.M[2] + ly * c.M[6] + lz * c.M[10];
        seen = cz > 0.0f && fabsf(cx) - radius < cz / c.fx * half_x && fabsf(cy) - radius < cz / c.fy * half_y;
    }
    const float g = grid[i];
    if (clear || (g < 0.0f) != !seen) grid[i] = seen ? 0.0f : -1.0f;
}

__global__ void dg_samples_kernel(uint32_t n, uint32_t step, float thresh, uint32_t G, const float* __restrict__ grid,
                                  Pcg32 rng, float* __restrict__ pos, uint32_t* __restrict__ idx_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t cells = G * G * G;
    rng.advance((uint64_t)i * 4u);
    (void)rng.next_float();  // cascade choice; one cascade here
    uint32_t idx = 0;
    for (uint32_t j = 0; j < 10; ++j) {
        idx = ((i + step * n) * 56924617u + j * 19349663u + 96925573u) % cells;
        if (grid[idx] > thresh) break;
    }
    const uint32_t x = compact3(idx), y = compact3(idx >> 1), z = compact3(idx >> 2);
    const float rx = rng.next_float(), ry = rng.next_float(), rz = rng.next_float();
    const float g = (float)G;
    pos[3 * i] = (((float)x + rx) / g - 0.5f) * 1.0f + 0.5f;
    pos[3 * i + 1] = (((float)y + ry) / g - 0.5f) * 1.0f + 0.5f;
    pos[3 * i + 2] = (((float)z + rz) / g - 0.5f) * 1.0f + 0.5f;
    idx_out[i] = idx;
}

// density MLP output 0 at a canonical position (the FMA chains of the plain
// inference kernel), optical thickness max-splatted into the sample's cell
template <int L, int X>
__global__ __launch_bounds__(128) void dg_splat_kernel(const float* __restrict__ params,
                                                       const float* __restrict__ table, GridParams gp,
                                                       const float* __restrict__ pos,
                                                       const uint32_t* __restrict__ idx, uint32_t n, float dt,
                                                       float bias, float* __restrict__ tmp) {
    constexpr int E = 2 * L + X;
    constexpr int W = 64;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 x = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    float enc[E];
#pragma unroll
    for (int l = 0; l < L; ++l) encode_level<2>(table, gp, (uint32_t)l, x, enc + 2 * l);
#pragma unroll
    for (int e = 2 * L; e < E; ++e) enc[e] = 1.0f;
    float h[W];
    dense<E, W, true>(params, enc, h);
    float logit;
    dense<W, 1, false>(params + W * E, h, &logit);
    const float thickness = expf(logit + bias) * dt;
    atomicMax(reinterpret_cast<uint32_t*>(tmp) + idx[i], __float_as_uint(thickness));
}

// NerfNetwork::density (nerf_network.h:385-414): the density MLP's 16
// outputs at canonical positions, extra dims 1.0 (extract_expressions_random
// writes 1.0, nerf_network.h:107-127); out [n][16], output 0 the density logit
template <int L, int X>
__global__ __launch_bounds__(128) void density_outputs_kernel(const float* __restrict__ params,
                                                              const float* __restrict__ table, GridParams gp,
                                                              const float* __restrict__ pos, uint32_t n,
                                                              float* __restrict__ out) {
    constexpr int E = 2 * L + X;
    constexpr int W = 64;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 x = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    float enc[E];
#pragma unroll
    for (int l = 0; l < L; ++l) encode_level<2>(table, gp, (uint32_t)l, x, enc + 2 * l);
#pragma unroll
    for (int e = 2 * L; e < E; ++e) enc[e] = 1.0f;
    float h[W];
    dense<E, W, true>(params, enc, h);
    float o[16];
    dense<W, 16, false>(params + W * E, h, o);
#pragma unroll
    for (int k = 0; k < 16; ++k) out[(size_t)i * 16 + k] = o[k];
}

__global__ void dg_ema_kernel(float* __restrict__ grid, const float* __restrict__ tmp, float decay, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float prev = grid[i];
    grid[i] = prev < 0.0f ? prev : fmaxf(prev * decay, tmp[i]);
}

// stage 1: block b sums max(grid, 0) / n over cells [b*kReduce, (b+1)*kReduce)
__global__ __launch_bounds__(kThreads) void dg_reduce_kernel(const float* __restrict__ grid, uint32_t n,
                                                              float* __restrict__ partial) {
    __shared__ float s[kThreads];
    const uint32_t b = blockIdx.x, t = threadIdx.x;
    const float inv = 1.0f / (float)n;
    float acc = 0.0f;
    for (uint32_t k = 0; k < kReduce / kThreads; ++k) {
        const uint32_t i = b * kReduce + k * kThreads + t;
        if (i < n) acc += fmaxf(grid[i], 0.0f) * inv;
    }
    s[t] = acc;
    __syncthreads();
    for (uint32_t o = kThreads / 2; o > 0; o >>= 1) {
        if (t < o) s[t] += s[t + o];
        __syncthreads();
    }
    if (t == 0) partial[b] = s[0];
}

// stage 2: one block sums the partials in a fixed order
__global__ __launch_bounds__(kThreads) void dg_mean_kernel(const float* __restrict__ partial, uint32_t nb,
                                                            float* __restrict__ mean) {
    __shared__ float s[kThreads];
    const uint32_t t = threadIdx.x;
    float acc = 0.0f;
    for (uint32_t i = t; i < nb; i += kThreads) acc += partial[i];
    s[t] = acc;
    __syncthreads();
    for (uint32_t o = kThreads / 2; o > 0; o >>= 1) {
        if (t < o) s[t] += s[t + o];
        __syncthreads();
    }
    if (t == 0) mean[0] = s[0];
}

__global__ void dg_bitfield_kernel(const float* __restrict__ grid, const float* __restrict__ mean,
                                   uint8_t* __restrict__ bits, uint32_t n_bytes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_bytes) return;
    const float thresh = fminf(kMinOpticalThickness, mean[0]);
    uint32_t v = 0;
    for (uint32_t j = 0; j < 8; ++j) v |= grid[8 * i + j] > thresh ? (1u << j) : 0u;
    bits[i] = (uint8_t)v;
}

bool dg_alloc(void* p, size_t bytes) { return hipMalloc(reinterpret_cast<void**>(p), bytes) == hipSuccess; }

}  // namespace

void free_density_grid(insta_ctx* ctx) {
    DensityGridState* d = ctx->density;
    if (!d) return;
    dev_free(d->grid); dev_free(d->tmp); dev_free(d->pos); dev_free(d->idx); dev_free(d->partial); dev_free(d->mean);
    if (d->h_mean) (void)hipHostFree(d->h_mean);
    delete d;
    ctx->density = nullptr;
}

insta_status update_density_grid(insta_ctx* ctx, uint32_t slot, uint32_t n_frames, const insta_density_grid_config& c,
                                 float* grid_out, float* mean_out, hipStream_t stream) {
    const uint32_t G = ctx->cfg.grid_size, cells = G * G * G;
    const uint32_t n = c.n_uniform + c.n_nonuniform;
    if (!ctx->density) {
        DensityGridState* d = new (std::nothrow) DensityGridState();
        if (!d) {
            set_error("host allocation failed");
            return INSTA_ERR_OOM;
        }
        ctx->density = d;
        const uint32_t nb = (cells + kReduce - 1) / kReduce;
        if (!dg_alloc(&d->grid, cells * 4) || !dg_alloc(&d->tmp, cells * 4) || !dg_alloc(&d->partial, nb * 4) ||
            !dg_alloc(&d->mean, 4) ||
            hipHostMalloc(reinterpret_cast<void**>(&d->h_mean), 4, hipHostMallocDefault) != hipSuccess) {
            free_density_grid(ctx);
            set_error("density grid allocation failed");
            return INSTA_ERR_OOM;
        }
    }
    DensityGridState* d = ctx->density;
    if (n > d->cap) {
        dev_free(d->pos);
        dev_free(d->idx);
        d->cap = 0;
        if (!dg_alloc(&d->pos, (size_t)n * 12) || !dg_alloc(&d->idx, (size_t)n * 4)) {
            set_error("density grid sample allocation failed");
            return INSTA_ERR_OOM;
        }
        d->cap = n;
    }
    const dim3 cgrid((cells + kThreads - 1) / kThreads), block(kThreads);
    if (d->step == 0) {
        d->rng = Pcg32::seeded(c.seed);
        hipLaunchKernelGGL(dg_fill_kernel, cgrid, block, 0, stream, d->grid, cells);
    }
    // cells no training camera sees, on the first update and whenever the
    // number of training images changed (the reference keys this to
    // training step 0, where its first update runs, testbed_nerf.cu:2795-2806)
    const TrainSetView tv = train_set_view(ctx);
    if (tv.n_images > 0 && (d->step == 0 || tv.n_images != d->marked_n)) {
        hipLaunchKernelGGL(dg_mark_kernel, cgrid, block, 0, stream, d->grid, G, tv.meta, tv.n_images,
                           (float)tv.width * 0.5f, (float)tv.height * 0.5f, d->step == 0 ? 1u : 0u);
        d->marked_n = tv.n_images;
    }
    hipLaunchKernelGGL(dg_fill_kernel, cgrid, block, 0, stream, d->tmp, cells);
    if (c.n_uniform)
        hipLaunchKernelGGL(dg_samples_kernel, dim3((c.n_uniform + kThreads - 1) / kThreads), block, 0, stream,
                           c.n_uniform, d->step, -0.01f, G, d->grid, d->rng, d->pos, d->idx);
    if (c.n_nonuniform)
        hipLaunchKernelGGL(dg_samples_kernel, dim3((c.n_nonuniform + kThreads - 1) / kThreads), block, 0, stream,
                           c.n_nonuniform, d->step, kMinOpticalThickness, G, d->grid, d->rng,
                           d->pos + 3 * (size_t)c.n_uniform, d->idx + c.n_uniform);
    INSTA_HIP(hipGetLastError());
    d->rng.advance(1ull << 32);
    // frames for the warp: a second stream, advanced per update
    Pcg32 wr = Pcg32::seeded(c.seed ^ 0x9e3779b9u);
    wr.advance((uint64_t)d->step << 32);
    insta_status st = launch_deform_positions(ctx, d->pos, n, n_frames, wr, stream);
    if (st != INSTA_OK) return st;
    const float* table = ctx->d_params + ctx->layout.n_mlp_params;
    const GridParams gp = grid_params(ctx);
    const dim3 sgrid((n + 127) / 128);
#define INSTA_SPLAT(L_, X_)                                                                                     \
    hipLaunchKernelGGL((dg_splat_kernel<L_, X_>), sgrid, dim3(128), 0, stream, ctx->d_params, table, gp, d->pos,    \
                       d->idx, n, ctx->cfg.dt, ctx->cfg.density_logit_bias, d->tmp)
    const bool x16 = ctx->cfg.n_extra_dims != 0;
    if (ctx->cfg.n_levels == 8) {
        if (x16) INSTA_SPLAT(8, 16);
        else INSTA_SPLAT(8, 0);
    } else {
        if (x16) INSTA_SPLAT(16, 16);
        else INSTA_SPLAT(16, 0);
    }
#undef INSTA_SPLAT
    INSTA_HIP(hipGetLastError());
    hipLaunchKernelGGL(dg_ema_kernel, cgrid, block, 0, stream, d->grid, d->tmp, c.decay, cells);
    const uint32_t nb = (cells + kReduce - 1) / kReduce;
    hipLaunchKernelGGL(dg_reduce_kernel, dim3(nb), block, 0, stream, d->grid, cells, d->partial);
    hipLaunchKernelGGL(dg_mean_kernel, dim3(1), block, 0, stream, d->partial, nb, d->mean);
    hipLaunchKernelGGL(dg_bitfield_kernel, dim3((cells / 8 + kThreads - 1) / kThreads), block, 0, stream, d->grid,
                       d->mean, ctx->grid(slot), cells / 8);
    INSTA_HIP(hipGetLastError());
    st = launch_grid_linear(ctx, slot, stream);
    if (st != INSTA_OK) return st;
    if (grid_out) INSTA_HIP(hipMemcpyAsync(grid_out, d->grid, cells * 4, hipMemcpyDeviceToDevice, stream));
    if (mean_out) {
        INSTA_HIP(hipMemcpyAsync(d->h_mean, d->mean, 4, hipMemcpyDeviceToHost, stream));
        INSTA_HIP(hipStreamSynchronize(stream));
        *mean_out = *d->h_mean;
    }
    ctx->grid_built[slot] = 1;
    ++d->step;
    return INSTA_OK;
}

insta_status launch_density_outputs(insta_ctx* ctx, const float* pos, uint32_t n, float* out, hipStream_t stream) {
    if (n == 0) return INSTA_OK;
    const GridParams gp = grid_params(ctx);
    const float* table = ctx->d_params + ctx->layout.n_mlp_params;
    const dim3 grid((n + 127) / 128), block(128);
    const bool x16 = ctx->cfg.n_extra_dims == 16;
    if (ctx->cfg.n_levels == 8)
        hipLaunchKernelGGL((x16 ? density_outputs_kernel<8, 16> : density_outputs_kernel<8, 0>), grid, block, 0, stream,
                           ctx->d_params, table, gp, pos, n, out);
    else
        hipLaunchKernelGGL((x16 ? density_outputs_kernel<16, 16> : density_outputs_kernel<16, 0>), grid, block, 0,
                           stream, ctx->d_params, table, gp, pos, n, out);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status density_export(insta_ctx* ctx, DensitySnapshot& out, hipStream_t stream) {
    out = DensitySnapshot{};
    DensityGridState* d = ctx->density;
    if (!d || d->step == 0) return INSTA_OK;  // never populated
    const uint32_t G = ctx->cfg.grid_size, cells = G * G * G;
    out.grid.resize(cells);
    // the grid was last written on the caller's stream: copy there (pageable
    // destination), without stalling other streams of the device
    INSTA_HIP(hipMemcpyAsync(out.grid.data(), d->grid, (size_t)cells * 4, hipMemcpyDeviceToHost, stream));
    INSTA_HIP(hipStreamSynchronize(stream));
    out.step = d->step;
    out.rng_state = d->rng.state;
    out.rng_inc = d->rng.inc;
    return INSTA_OK;
}

insta_status density_import(insta_ctx* ctx, const DensitySnapshot& in, uint32_t slot, hipStream_t stream) {
    const uint32_t G = ctx->cfg.grid_size, cells = G * G * G;
    if (in.grid.empty()) {
        if (ctx->density) ctx->density->step = 0;
        return INSTA_OK;
    }
    if (in.grid.size() != cells) {
        set_error("density grid: size does not match the context's grid");
        return INSTA_ERR_INVALID;
    }
    if (!ctx->density) {
        DensityGridState* d = new (std::nothrow) DensityGridState();
        if (!d) {
            set_error("host allocation failed");
            return INSTA_ERR_OOM;
        }
        ctx->density = d;
        const uint32_t nb = (cells + kReduce - 1) / kReduce;
        if (!dg_alloc(&d->grid, cells * 4) || !dg_alloc(&d->tmp, cells * 4) || !dg_alloc(&d->partial, nb * 4) ||
            !dg_alloc(&d->mean, 4) ||
            hipHostMalloc(reinterpret_cast<void**>(&d->h_mean), 4, hipHostMallocDefault) != hipSuccess) {
            free_density_grid(ctx);
            set_error("density grid allocation failed");
            return INSTA_ERR_OOM;
        }
    }
    DensityGridState* d = ctx->density;
    INSTA_HIP(hipMemcpyAsync(d->grid, in.grid.data(), (size_t)cells * 4, hipMemcpyHostToDevice, stream));
    d->step = in.step ? in.step : 1;
    d->rng.state = in.rng_state;
    d->rng.inc = in.rng_inc;
    // mean and bitfield as after an update (update_density_grid_mean_and_bitfield)
    const dim3 block(kThreads);
    const uint32_t nb = (cells + kReduce - 1) / kReduce;
    hipLaunchKernelGGL(dg_reduce_kernel, dim3(nb), block, 0, stream, d->grid, cells, d->partial);
    hipLaunchKernelGGL(dg_mean_kernel, dim3(1), block, 0, stream, d->partial, nb, d->mean);
    hipLaunchKernelGGL(dg_bitfield_kernel, dim3((cells / 8 + kThreads - 1) / kThreads), block, 0, stream, d->grid,
                       d->mean, ctx->grid(slot), cells / 8);
    INSTA_HIP(hipGetLastError());
    insta_status st = launch_grid_linear(ctx, slot, stream);
    if (st != INSTA_OK) return st;
    INSTA_HIP(hipStreamSynchronize(stream));  // pageable source
    ctx->grid_built[slot] = 1;
    return INSTA_OK;
}

insta_status reset_density_grid(insta_ctx* ctx) {
    if (ctx->density) ctx->density->step = 0;
    return INSTA_OK;
}

}  // namespace insta
