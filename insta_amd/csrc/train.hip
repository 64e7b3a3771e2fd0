// train.hip -- one training step of the avatar network on the device.
//
// Restates Testbed::train_nerf_step (instant-ngp/src/testbed_nerf.cu:
// 3176-3400) for the INSTA path, as this sequence of kernels on one stream:
//   rays      generate_training_samples_nerf (1105-1285): per ray a training
//             image ((i + rays_total) * n_images / n_rays, 1082-1103) and a
//             PCG32 stream advanced by 8 draws per ray (1143): pixel xy, a
//             draw the reference spends on motion blur, the start jitter;
//             march through the image frame's occupancy grid from
//             max(t_box, 0) + dt * jitter, at most max_steps samples; a ray
//             whose sample range would cross the cap is dropped (1239-1242)
//   deform    posed -> canonical with the frame of each ray (core.cu:442-526)
//   infer     the context's inference network (MFMA) on every marched sample
//   loss      compute_loss_kernel_train_nerf (1305-1643): product
//             transmittance T *= 1 - alpha until T < 1e-4, random background
//             (the same PCG32 stream: draws 2-4 of the ray), Huber(0.1) / 5
//             per channel, dL/d(rgb logit) = w * dL/drgb * s(1 - s),
//             dL/d(density logit) = sigma * dt * dL/drgb . (T rgb - suffix),
//             scaled by loss_scale / n_rays; the composited samples are
//             compacted into a batch of at most target_samples (1543-1560)
//   forward   on the compacted batch: hash grid (FP32, the plain encode)
//             and the MLPs on FP32 MFMA, keeping every activation the
//             backward needs, feature-major [feature][sample]
//   backward  both MLPs on FP32 MFMA over the compacted batch (ReLU masks
//             from the stored activations) into dL/dencoding, then scattered
//             into the hash tables, one thread per sample: FP32 atomics, or
//             packed FP16 atomics (one global_atomic_pk_add_f16 per corner
//             and level, F = 2)
//   wgrad     weight gradients dW = sum_n delta_n x_n^T as per-chunk partial
//             products on FP32 MFMA (samples as the K dimension, operands
//             staged through LDS), then a fixed-order sum over chunks:
//             deterministic MLP gradients
//   adam      Kingma & Ba with bias correction over all parameters, after
//             removing the loss scale
// Extra dims (16 per frame) enter the density MLP as 16 more inputs, 1.0 or
// the ray's frame values by the selector the deformation writes; their
// columns of W0 get gradients like the encoding's, the values themselves
// none (latent-code optimisation is off in the reference configuration).
// Training images hold sRGB colour with straight alpha, as the reference's
// 8-bit images do.  Declared simplifications against the reference: no
// density/colour regularisers, no error map, rays never masked out.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include <type_traits>

#include "context.hpp"
#include "encoding.hpp"
#include "march.hpp"

namespace insta {

enum : uint32_t {
    kTRays = 0,      // rays kept (ray list length)
    kTSamples = 1,   // samples reserved
    kTLimit = 2,     // first sample of the ray that crossed the cap
    kTValid = 3,     // samples marched into the buffers
    kTUsed = 4,      // samples composited before termination
    kTLoss = 5,      // float bits: sum of per-ray loss / n_rays
    kTCompactRes = 6,  // compacted samples reserved
    kTCompactN = 7,    // compacted samples kept (<= compact capacity)
    kTRound0 = 8,      // samples evaluated in each window (kRounds counters)
    kTCounters = 12
};

// Evaluation windows over each ray's marched samples: samples [kWin[w],
// kWin[w + 1]) of the rays still alive are deformed and evaluated in round
// w, then each ray's transmittance advances over them (the loss kernel's
// arithmetic); a ray whose transmittance fell below 1e-4, or whose samples
// ran out, takes no further round.  The loss reads exactly the samples the
// rounds evaluated, so the step computes what one pass over every marched
// sample would.  Once trained, most rays stop within a sample or two of the
// surface, so the first window holds one sample (a first window of 4 or 8
// measured 10-90 % slower rounds).
constexpr uint32_t kRounds = 4;
constexpr uint32_t kWin[kRounds + 1] = {0u, 1u, 4u, 16u, 0xffffffffu};

constexpr int kW = 64;
constexpr int kDOut = 16;
constexpr int kCin = 32;             // density outputs + 16 SH
constexpr uint32_t kMlpThreads = 256;  // backward: one weight copy in LDS per block
constexpr uint32_t kChunk = 1024;    // samples per weight-gradient partial
constexpr uint32_t kRaySeedStride = 8;
// backward: wave-level run sums on levels where one marching step covers at
// most this fraction of a cell (runs of about two or more samples per entry)
constexpr float kRunMaxStepCells = 0.5f;

// activation rows (feature-major, row stride = sample capacity)
template <int E>
struct Act {
    static constexpr int ENC = 0, H1 = E, DENS = E + kW, SH = E + kW + kDOut, H2 = E + kW + kCin,
                         H3 = H2 + kW, N = H3 + kW;
};
// delta rows
struct Dact {
    static constexpr int DRGB = 0, DH3 = 4, DH2 = DH3 + kW, DDENS = DH2 + kW, DH1 = DDENS + kDOut, N = DH1 + kW;
};
// parameter offsets (density MLP, colour MLP), row-major [out][in]
template <int E>
struct Off {
    static constexpr uint32_t W0 = 0, W1 = W0 + kW * E, W2 = W1 + kDOut * kW, W3 = W2 + kW * kCin, W4 = W3 + kW * kW,
                              N = W4 + 3 * kW;
};

// XCDs of the device: workgroups of one XCD share its L2
constexpr uint32_t kXcds = 8;

struct TrainState {
    uint32_t n_images = 0, width = 0, height = 0;
    float4* images = nullptr;
    TrainImage* meta = nullptr;
    // supervision of the training set (nullptr: none): mask id and
    // camera-axis depth per pixel, [n][h][w]
    uint8_t* seg = nullptr;
    float* depth = nullptr;
    std::vector<int32_t> frames;
    // workspace
    uint32_t ray_cap = 0, sample_cap = 0, compact_cap = 0;
    float4* ray_d = nullptr;
    float4* ray_o = nullptr;       // origin, w = length of the unnormalised direction
    int32_t* ray_frame = nullptr;
    uint32_t *ray_img = nullptr, *ray_pix = nullptr, *ray_src = nullptr, *ray_base = nullptr, *ray_cnt = nullptr;
    float4 *s_cpos = nullptr, *s_cdir = nullptr, *s_out = nullptr, *s_dout = nullptr;
    // evaluation rounds: per ray transmittance, alive flag and the first slot
    // of its samples in the round buffers (r_*: one window's samples, packed)
    float* ray_T = nullptr;
    uint32_t *ray_alive = nullptr, *ray_gbase = nullptr;
    float4 *r_pos = nullptr, *r_cpos = nullptr, *r_cdir = nullptr, *r_out = nullptr;
    int32_t* s_hit = nullptr;
    uint32_t* c_idx = nullptr;     // compacted batch: sample index
    float4* c_dout = nullptr;      // compacted batch: dL/d(rgb logits, density logit)
    float *act = nullptr, *dact = nullptr, *part = nullptr;
    float* t_buf = nullptr;        // march t values, step-major [max_steps][rays of the step]: the rounds and the loss read them
    // optimizer
    float *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    // hash-table gradients during the backward: one copy per XCD (FP32 or
    // packed FP16), summed into grad afterwards
    float* grad_x = nullptr;
    __half2* grad_xh = nullptr;
    float* denc = nullptr;         // dL/dencoding of the batch [32][batch cap], for the LDS scatter
    // LDS-scattered levels (a prefix of the levels, entries [0, lds_end)):
    // one partial table per chunk of the batch, [kScatterMaxChunks][2 lds_end]
    float* grad_p = nullptr;
    uint32_t lds_levels = 0, lds_end = 0;
    uint32_t* counters = nullptr;
    uint32_t* h_counters = nullptr;
    uint64_t rays_total = 0;
    uint32_t rays_per_batch = 1u << 12;
    // two counters, as the reference keeps them: the training step
    // (Testbed::m_training_step, which times the density-grid updates and
    // resumes at the saved value + 1 on snapshot load, testbed.cu:3011) and
    // Adam's own step (bias correction; restarts with fresh moments)
    uint32_t training_step = 0;
    uint32_t adam_step = 0;
    // Ema optimizer: running average of the weights and its debiased copy
    float *ema = nullptr, *ema_infer = nullptr;
    uint32_t ema_step = 0;
    bool have_grad = false;
    // network plugin: per-sample extra dims of the last forward ([cap][X]) and
    // its batch size (0: no stored forward; a training step overwrites it)
    float* net_extra = nullptr;
    uint32_t net_extra_cap = 0;
    uint32_t net_n = 0;
};

namespace {

__global__ void train_clear_kernel(uint32_t* counters, uint32_t limit) {
    const uint32_t i = threadIdx.x;
    if (i < kTCounters) counters[i] = i == kTLimit ? limit : 0u;
}

__global__ void fill_zero_kernel(uint32_t* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void train_count_kernel(uint32_t* counters) {
    if (threadIdx.x == 0) counters[kTValid] = min(counters[kTSamples], counters[kTLimit]);
}

__global__ void train_compact_count_kernel(uint32_t* counters, uint32_t cap) {
    if (threadIdx.x == 0) counters[kTCompactN] = min(counters[kTCompactRes], cap);
}

struct RayArgs {
    RenderArgs box;               // render box, G, dt, max_steps (grid set per ray)
    const TrainImage* meta;
    const uint8_t* grids_lin;     // [slot][G^3/8]
    size_t grid_bytes;
    uint32_t n_images, width, height, max_samples;
    int32_t fixed_slot;           // >= -1: every ray marches this slot's grid
};

INSTA_DEV bool train_ray(const RayArgs& a, uint32_t i, uint32_t n_rays, uint64_t rays_total, Pcg32 rng,
                         uint32_t& img, uint32_t& pix, V3& o, V3& d, float& dn, float& t0, RenderArgs& ra) {
    img = (uint32_t)((((uint64_t)i + rays_total) * a.n_images / n_rays) % a.n_images);
    rng.advance((uint64_t)i * kRaySeedStride);
    const float x = rng.next_float();
    const float y = rng.next_float();
    (void)rng.next_float();  // motion-blur time in the reference
    const float jitter = rng.next_float();
    const uint32_t px = min((uint32_t)(x * (float)a.width), a.width - 1u);
    const uint32_t py = min((uint32_t)(y * (float)a.height), a.height - 1u);
    pix = px + py * a.width;
    const TrainImage& c = a.meta[img];
    const V3 d0 = v3((x - c.cx) * (float)a.width / c.fx, (y - c.cy) * (float)a.height / c.fy, 1.0f);
    d = v3(c.M[0] * d0.x + c.M[1] * d0.y + c.M[2] * d0.z, c.M[4] * d0.x + c.M[5] * d0.y + c.M[6] * d0.z,
           c.M[8] * d0.x + c.M[9] * d0.y + c.M[10] * d0.z);
    dn = norm(d);
    d = v3(d.x / dn, d.y / dn, d.z / dn);
    o = v3(c.M[3], c.M[7], c.M[11]);
    ra = a.box;
    ra.grid = a.grids_lin + (size_t)((a.fixed_slot >= -1 ? a.fixed_slot : c.frame) + 1) * a.grid_bytes;
    const float tb = box_ray_tmin(ra, o, d);
    if (tb == 3.40282347e38f) return false;
    t0 = fmaxf(tb, 0.0f) + a.box.dt * jitter;
    return true;
}

__global__ __launch_bounds__(256) void train_rays_kernel(RayArgs a, uint32_t n_rays, uint64_t rays_total,
                                                         Pcg32 rng, float4* __restrict__ ray_d,
                                                         float4* __restrict__ ray_o,
                                                         int32_t* __restrict__ ray_frame,
                                                         uint32_t* __restrict__ ray_img,
                                                         uint32_t* __restrict__ ray_pix,
                                                         uint32_t* __restrict__ ray_src,
                                                         uint32_t* __restrict__ ray_base,
                                                         uint32_t* __restrict__ ray_cnt,
                                                         float* __restrict__ t_buf,
                                                         float* __restrict__ ray_T,
                                                         uint32_t* __restrict__ ray_alive,
                                                         uint32_t* __restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t img = 0, pix = 0, count = 0;
    V3 o = v3(0, 0, 0), d = v3(0, 0, 1), idir = v3(0, 0, 1);
    float t0 = 0.0f, dn = 1.0f;
    RenderArgs ra = a.box;
    if (i < n_rays && train_ray(a, i, n_rays, rays_total, rng, img, pix, o, d, dn, t0, ra)) {
        idir = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        // one iteration per sample or empty-voxel skip; sample t values to
        // step-major scratch (t_buf[j * n_rays + i])
        float t = t0;
        while (count < a.box.max_steps) {
            const V3 pos = o + d * t;
            if (!box_contains(ra, pos)) break;
            if (occupied(ra.grid, ra.G, pos)) {
                t_buf[(size_t)count * n_rays + i] = t;
                ++count;
                t += a.box.dt;
            } else {
                t = advance_to_next_voxel(t, a.box.dt, pos, d, idir, ra.G);
            }
        }
    }
    const uint32_t base = wave_reserve(&counters[kTSamples], count);
    const bool fits = base + count <= a.max_samples;
    if (count > 0 && !fits && base < a.max_samples) atomicMin(&counters[kTLimit], base);
    const bool keep = count > 0 && fits;
    const uint32_t r = wave_append(&counters[kTRays], keep);
    if (!keep) return;
    ray_d[r] = make_float4(d.x, d.y, d.z, 0.0f);
    ray_o[r] = make_float4(o.x, o.y, o.z, dn);
    ray_frame[r] = a.meta[img].frame;
    ray_img[r] = img;
    ray_pix[r] = pix;
    ray_src[r] = i;
    ray_base[r] = base;
    ray_cnt[r] = count;
    ray_T[r] = 1.0f;
    ray_alive[r] = 1u;
}

template <int NIN, int NOUT>
INSTA_DEV void store_rows(float* __restrict__ act, uint32_t ld, int row0, uint32_t i, const float* v) {
#pragma unroll
    for (int k = 0; k < NOUT; ++k) act[(size_t)(row0 + k) * ld + i] = v[k];
}

// Forward with stored activations over the compacted batch, on
// v_mfma_f32_16x16x4f32 (the layout of the FP32 inference kernel): phase A,
// one thread per sample, encodes (FP32, the scalar encode's arithmetic) into
// LDS and stores the encoding, extra-dims and SH rows; phase B runs each
// wave's 64 samples, 16 columns at a time, through the density MLP and the
// colour MLP's hidden layers with the weights as A operands, storing every
// layer's rows from the accumulators (lane l holds rows 4(l/16)..+3 of
// column l%16, contiguous along the batch).  The training step needs no
// output layer (dL/doutput comes from the loss kernel); the network plugin's
// forward asks for it through `out` (rgb logits and the density logit per
// batch entry).  Extra dims come from the frame table by the selector, or
// per sample from x_extra [n][X] when given.
typedef float tf_f4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kFwdTile = 256;
template <int L, int X>
__global__ __launch_bounds__(256) void train_forward_kernel(const float* __restrict__ params,
                                                          const float* __restrict__ table, GridParams gp,
                                                          const float* __restrict__ frame_extra, uint32_t n_frames,
                                                          const uint32_t* __restrict__ n_dev,
                                                          const uint32_t* __restrict__ c_idx,
                                                          const float4* __restrict__ cpos,
                                                          const float4* __restrict__ cdir,
                                                          const float* __restrict__ x_extra,
                                                          float* __restrict__ act, uint32_t ld,
                                                          float4* __restrict__ out) {
    constexpr int E = 2 * L, DIN = E + X, KB0 = DIN / 4, S = DIN + 1;
    static_assert(DIN % 4 == 0, "K blocks of 4");
    using A = Act<DIN>;
    using O = Off<DIN>;
    __shared__ float s_x[kFwdTile * S];
    const uint32_t n = *n_dev;
    const uint32_t base = blockIdx.x * kFwdTile;
    if (base >= n) return;  // whole block past the batch
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t m = lane & 15u, g = lane >> 4;
    {
        const uint32_t i = base + tid;
        float* row = s_x + tid * S;
        if (i < n) {
            const uint32_t smp = c_idx[i];
            const float4 p4 = cpos[smp], d4 = cdir[smp];
            const V3 x = v3(p4.x, p4.y, p4.z);
            float enc[DIN];
#pragma unroll
            for (int l = 0; l < L; ++l) encode_level<2>(table, gp, (uint32_t)l, x, enc + 2 * l);
            if constexpr (X > 0) {
                const uint32_t sel = __float_as_uint(d4.w);
                const float* fx = x_extra ? x_extra + (size_t)smp * X
                                          : (sel != kExtraOnes && sel <= n_frames ? frame_extra + (size_t)(sel - 1u) * X
                                                                                  : nullptr);
#pragma unroll
                for (int e = 0; e < X; ++e) enc[E + e] = fx ? fx[e] : 1.0f;
            }
#pragma unroll
            for (int k = 0; k < DIN; ++k) row[k] = enc[k];
            store_rows<DIN, DIN>(act, ld, A::ENC, i, enc);
            float sh[16];
            sh16(d4.x * 2.0f - 1.0f, d4.y * 2.0f - 1.0f, d4.z * 2.0f - 1.0f, sh);
            store_rows<16, 16>(act, ld, A::SH, i, sh);
        } else {
#pragma unroll
            for (int k = 0; k < DIN; ++k) row[k] = 0.0f;
        }
    }
    __syncthreads();
    const float* W0 = params + O::W0;
    const float* W1 = params + O::W1;
    const float* W2 = params + O::W2;
    const float* W3 = params + O::W3;
    for (uint32_t nb = 0; nb < 4; ++nb) {
        const uint32_t s = wave * 64 + nb * 16 + m;  // column (sample) of this lane
        const uint32_t i = base + s;
        const bool live = i < n;
        const float* xr = s_x + s * S;
        // SH of the column's sample; lane group g feeds K rows 4g..4g+3
        float shg[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (live) {
            const float4 d4 = cdir[c_idx[i]];
            float sh[16];
            sh16(d4.x * 2.0f - 1.0f, d4.y * 2.0f - 1.0f, d4.z * 2.0f - 1.0f, sh);
#pragma unroll
            for (int b = 0; b < 4; ++b) shg[b] = g == 0 ? sh[b] : (g == 1 ? sh[4 + b] : (g == 2 ? sh[8 + b] : sh[12 + b]));
        }
        // density layer 0: H1 = relu(W0 x)
        tf_f4 h[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            tf_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < KB0; ++kb)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W0[(16 * t + m) * DIN + 4 * kb + g], xr[4 * kb + g], acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                h[t][q] = fmaxf(acc[q], 0.0f);
                if (live) act[(size_t)(A::H1 + 16 * t + 4 * g + q) * ld + i] = h[t][q];
            }
        }
        // density output: 16 rows, K = H1 in the order 16t + 4g + b
        tf_f4 dn = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                dn = __builtin_amdgcn_mfma_f32_16x16x4f32(W1[m * kW + 16 * t + 4 * g + b], h[t][b], dn, 0, 0, 0);
        if (live)
#pragma unroll
            for (int q = 0; q < 4; ++q) act[(size_t)(A::DENS + 4 * g + q) * ld + i] = dn[q];
        // colour layer 0: K = density outputs (4g + b) then SH (16 + 4g + b)
        tf_f4 r[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            tf_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W2[(16 * t + m) * kCin + 4 * g + b], dn[b], acc, 0, 0, 0);
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W2[(16 * t + m) * kCin + 16 + 4 * g + b], shg[b], acc, 0, 0,
                                                           0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                r[t][q] = fmaxf(acc[q], 0.0f);
                if (live) act[(size_t)(A::H2 + 16 * t + 4 * g + q) * ld + i] = r[t][q];
            }
        }
        // colour layer 1
        tf_f4 h3[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            tf_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W3[(16 * t + m) * kW + 16 * s2 + 4 * g + b], r[s2][b],
                                                               acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                h3[t][q] = fmaxf(acc[q], 0.0f);
                if (live) act[(size_t)(A::H3 + 16 * t + 4 * g + q) * ld + i] = h3[t][q];
            }
        }
        if (out) {
            // colour output (3 rows of W4, padded to 16): lane group 0 holds
            // rows 0..3 of its column; the density logit is density output 0
            const float* W4 = params + O::W4;
            tf_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(m < 3 ? W4[m * kW + 16 * s2 + 4 * g + b] : 0.0f,
                                                               h3[s2][b], acc, 0, 0, 0);
            if (live && g == 0) out[i] = make_float4(acc[0], acc[1], acc[2], dn[0]);
        }
    }
}

struct LossArgs {
    const float4* images;
    const uint8_t* seg;   // mask ids (nullptr: no masks)
    const float* depth;   // camera-axis depth, 0 = none (nullptr: no depth images)
    uint32_t width, height;
    float dt, density_bias, loss_scale_per_ray, inv_rays;
    float seg_weight, depth_weight;
};

INSTA_DEV float sigmoid(float x) { return 1.0f / (1.0f + expf(-x)); }

// one sample's transmittance step, exactly as the loss kernel composites
INSTA_DEV float transmit(float T, float density_logit, float density_bias, float dt) {
    const float sigma = expf(density_logit + density_bias);
    const float alpha = 1.0f - expf(-sigma * dt);
    return T * (1.0f - alpha);
}

// Round `w` gather: the window's samples of every alive ray, packed into the
// round buffers (each ray's samples contiguous from ray_gbase).
__global__ __launch_bounds__(256) void train_gather_kernel(uint32_t lo, uint32_t hi,
                                                           const uint32_t* __restrict__ ray_base,
                                                           const uint32_t* __restrict__ ray_cnt,
                                                           const uint32_t* __restrict__ ray_alive,
                                                           uint32_t* __restrict__ ray_gbase,
                                                           const float4* __restrict__ ray_o,
                                                           const float4* __restrict__ ray_d,
                                                           const uint32_t* __restrict__ ray_src,
                                                           const float* __restrict__ t_buf, uint32_t t_stride,
                                                           float4* __restrict__ r_pos,
                                                           uint32_t* __restrict__ counters, uint32_t round_slot) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = 0;
    if (r < counters[kTRays] && ray_alive[r]) {
        const uint32_t cnt = ray_cnt[r];
        n = cnt > lo ? min(cnt, hi) - lo : 0u;
    }
    const uint32_t g = wave_reserve(&counters[round_slot], n);
    if (n == 0) return;
    ray_gbase[r] = g;
    // posed positions from the march's t values (the march's own arithmetic)
    const float4 o4 = ray_o[r], d4 = ray_d[r];
    const V3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
    const float* __restrict__ tr = t_buf + ray_src[r];
    for (uint32_t k = 0; k < n; ++k) {
        const V3 pos = o + d * tr[(size_t)(lo + k) * t_stride];
        r_pos[g + k] = make_float4(pos.x, pos.y, pos.z, __uint_as_float(r));
    }
}

// Round `w` scatter and transmittance: each alive ray's evaluated samples go
// back to their marched slots, and its transmittance advances over them; the
// ray stays alive only while it is >= 1e-4 with samples left past the window.
__global__ __launch_bounds__(256) void train_round_kernel(uint32_t lo, uint32_t hi, float density_bias, float dt,
                                                          const uint32_t* __restrict__ ray_base,
                                                          const uint32_t* __restrict__ ray_cnt,
                                                          uint32_t* __restrict__ ray_alive,
                                                          const uint32_t* __restrict__ ray_gbase,
                                                          float* __restrict__ ray_T,
                                                          const float4* __restrict__ r_cpos,
                                                          const float4* __restrict__ r_cdir,
                                                          const float4* __restrict__ r_out,
                                                          float4* __restrict__ s_cpos, float4* __restrict__ s_cdir,
                                                          float4* __restrict__ s_out,
                                                          const uint32_t* __restrict__ counters) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= counters[kTRays] || !ray_alive[r]) return;
    const uint32_t cnt = ray_cnt[r];
    const uint32_t n = cnt > lo ? min(cnt, hi) - lo : 0u;
    const uint32_t base = ray_base[r] + lo, g = ray_gbase[r];
    float T = ray_T[r];
    bool alive = n > 0;
    for (uint32_t k = 0; k < n; ++k) {
        const float4 o = r_out[g + k];
        s_cpos[base + k] = r_cpos[g + k];
        s_cdir[base + k] = r_cdir[g + k];
        s_out[base + k] = o;
        if (alive) {
            T = transmit(T, o.w, density_bias, dt);
            if (T < 1e-4f) alive = false;
        }
    }
    ray_T[r] = T;
    ray_alive[r] = alive && cnt > hi ? 1u : 0u;
}
// sRGB transfer functions with the reference's constants (common_device.cuh:31-61)
INSTA_DEV float srgb_to_linear(float c) { return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f); }
INSTA_DEV float linear_to_srgb(float c) { return c < 0.0031308f ? 12.92f * c : 1.055f * powf(c, 0.41666f) - 0.055f; }

__global__ __launch_bounds__(256) void train_loss_kernel(LossArgs a, Pcg32 rng0, const uint32_t* __restrict__ ray_src,
                                                         const uint32_t* __restrict__ ray_img,
                                                         const uint32_t* __restrict__ ray_pix,
                                                         const uint32_t* __restrict__ ray_base,
                                                         const uint32_t* __restrict__ ray_cnt,
                                                         const float4* __restrict__ ray_o,
                                                         const float4* __restrict__ ray_d,
                                                         const float* __restrict__ t_buf, uint32_t t_stride,
                                                         const float4* __restrict__ s_out,
                                                         uint32_t* __restrict__ c_idx, float4* __restrict__ c_dout,
                                                         uint32_t compact_cap, uint32_t* __restrict__ counters) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = r < counters[kTRays];
    float ray_loss = 0.0f;
    uint32_t used = 0, base = 0;
    float acc[3] = {0.0f, 0.0f, 0.0f}, grad[3] = {0.0f, 0.0f, 0.0f};
    // depth supervision: the composited depth and dL/dD (0: none for this ray)
    float depth_ray = 0.0f, dgrad = 0.0f;
    V3 org = v3(0.0f, 0.0f, 0.0f), dir = v3(0.0f, 0.0f, 1.0f);
    const float* __restrict__ tr = t_buf;  // this ray's march t values, stride t_stride
    if (valid) {
        base = ray_base[r];
        tr = t_buf + ray_src[r];
        const uint32_t cnt = ray_cnt[r];
        Pcg32 rng = rng0;
        rng.advance((uint64_t)ray_src[r] * kRaySeedStride);
        (void)rng.next_float();
        (void)rng.next_float();
        // linear colour space with sRGB training (testbed_nerf.cu:1419-1443):
        // the random background goes to linear, is blended behind the
        // premultiplied linear texel (read_rgba of an 8-bit sRGB image,
        // common_device.cuh:658-671), then target and background return to sRGB
        const size_t texel = (size_t)ray_img[r] * a.width * a.height + ray_pix[r];
        const float4 tex = a.images[texel];
        const float texc[3] = {tex.x, tex.y, tex.z};
        float bg[3], target[3];
        for (int c = 0; c < 3; ++c) {
            const float bg_lin = srgb_to_linear(rng.next_float());
            bg[c] = linear_to_srgb(bg_lin);
            target[c] = linear_to_srgb(srgb_to_linear(texc[c]) * tex.w + (1.0f - tex.w) * bg_lin);
        }
        // segmentation and depth supervision (testbed_nerf.cu:1476-1497): mask
        // ids 120/130/140 weight the colour gradient; ids <= 90 or > 250 and
        // near-transparent texels get no depth term; the target is the depth
        // texel times the unnormalised direction's length (0 texel: none)
        float seg_w = 1.0f, lam = a.depth_weight;
        if (a.seg) {
            const uint32_t id = a.seg[texel];
            if (id == 120u || id == 130u || id == 140u) seg_w = a.seg_weight;
            if (id <= 90u || id > 250u) lam = 0.0f;
            if (tex.w < 0.001f) lam = 0.0f;
        }
        const float4 ro = ray_o[r], rd = ray_d[r];
        org = v3(ro.x, ro.y, ro.z);
        dir = v3(rd.x, rd.y, rd.z);
        const float target_depth = ro.w * ((lam > 0.0f && a.depth) ? a.depth[texel] : -1.0f);
        const bool supervise_depth = target_depth > 0.0f;
        float T = 1.0f;
        for (; used < cnt; ++used) {
            if (T < 1e-4f) break;
            const float4 o = s_out[base + used];
            const float sigma = expf(o.w + a.density_bias);
            const float alpha = 1.0f - expf(-sigma * a.dt);
            const float w = alpha * T;
            acc[0] = acc[0] + w * sigmoid(o.x);
            acc[1] = acc[1] + w * sigmoid(o.y);
            acc[2] = acc[2] + w * sigmoid(o.z);
            if (supervise_depth) {
                const V3 p = org + dir * tr[(size_t)used * t_stride];
                depth_ray = depth_ray + w * norm(p - org);
            }
            T = T * (1.0f - alpha);  // = transmit(T, o.w, ...), which the rounds use
        }
        // L1 depth loss (depth_loss_type's default, testbed_nerf.cu:133-139)
        if (supervise_depth) dgrad = copysignf(1.0f, depth_ray - target_depth) * lam;
        if (used == cnt) {
            for (int c = 0; c < 3; ++c) acc[c] = acc[c] + T * bg[c];
        }
        // Huber(0.1) / 5 per channel (testbed_nerf.cu:141-157, 1299)
        for (int c = 0; c < 3; ++c) {
            const float diff = acc[c] - target[c];
            const float ad = fabsf(diff);
            const float l = ad > 0.1f ? ad - 0.05f : (0.5f / 0.1f) * diff * diff;
            grad[c] = (ad > 0.1f ? (diff > 0.0f ? 1.0f : -1.0f) : diff / 0.1f) / 5.0f * seg_w;
            ray_loss += l / 5.0f;
        }
        ray_loss = ray_loss / 3.0f * a.inv_rays;
    }
    // compaction of the composited samples into the training batch
    const uint32_t cb = wave_reserve(&counters[kTCompactRes], used);
    const uint32_t n_c = cb < compact_cap ? min(used, compact_cap - cb) : 0u;
    if (valid) {
        const float ls = a.loss_scale_per_ray;
        float T = 1.0f;
        float acc2[3] = {0.0f, 0.0f, 0.0f}, depth2 = 0.0f;
        for (uint32_t j = 0; j < n_c; ++j) {
            const float4 o = s_out[base + j];
            const float sg[3] = {sigmoid(o.x), sigmoid(o.y), sigmoid(o.z)};
            const float sigma = expf(o.w + a.density_bias);
            const float alpha = 1.0f - expf(-sigma * a.dt);
            const float w = alpha * T;
            for (int c = 0; c < 3; ++c) acc2[c] = acc2[c] + w * sg[c];
            T = T * (1.0f - alpha);
            float dsig = 0.0f;
            float drgb[3];
            for (int c = 0; c < 3; ++c) {
                drgb[c] = ls * (w * grad[c] * (sg[c] * (1.0f - sg[c])));
                dsig = dsig + grad[c] * (T * sg[c] - (acc[c] - acc2[c]));
            }
            if (dgrad != 0.0f) {
                // dD/dsigma_j / dt = T_{j+1} depth_j - (depth suffix after j)
                const V3 p = org + dir * tr[(size_t)j * t_stride];
                const float dj = norm(p - org);
                depth2 = depth2 + w * dj;
                dsig = dsig + dgrad * (T * dj - (depth_ray - depth2));
            }
            c_idx[cb + j] = base + j;
            c_dout[cb + j] = make_float4(drgb[0], drgb[1], drgb[2], ls * (sigma * (a.dt * dsig)));
        }
    }
    // statistics
    uint32_t u = used;
    float lsum = ray_loss;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        u += __shfl_xor(u, off);
        lsum += __shfl_xor(lsum, off);
    }
    if ((threadIdx.x & 63u) == 0) {
        if (u) atomicAdd(&counters[kTUsed], u);
        if (lsum != 0.0f) atomicAdd(reinterpret_cast<float*>(&counters[kTLoss]), lsum);
    }
}

// The XCD (accelerator complex die) this wave runs on, 0 .. kXcds - 1: the
// hardware register XCC_ID (s_getreg_b32, register 20).
INSTA_DEV uint32_t xcd_id() { return (uint32_t)__builtin_amdgcn_s_getreg((31 << 11) | 20) % kXcds; }

// Table gradients into grad_tab (FP32): entries below lds_end are the sum of
// the LDS scatter's chunk partials (chunk order), the others the sum of the
// per-XCD copies (a fixed order, the FP32 ones then the FP16 ones, each set
// when in use; the copies are cleared before each backward).
__global__ void grad_copies_reduce_kernel(const float* __restrict__ grad_x, const __half2* __restrict__ grad_xh,
                                          uint64_t entries, uint32_t use_x, uint32_t use_xh,
                                          const float* __restrict__ grad_p, uint32_t lds_end, uint32_t chunks,
                                          float* __restrict__ grad_tab, uint32_t accumulate) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < entries;
         e += (uint64_t)gridDim.x * blockDim.x) {
        float a = 0.0f, b = 0.0f;
        if (e < lds_end) {
            for (uint32_t c = 0; c < chunks; ++c) {
                const float2 v = *reinterpret_cast<const float2*>(grad_p + 2 * ((uint64_t)c * lds_end + e));
                a = a + v.x;
                b = b + v.y;
            }
        } else if (use_x)
#pragma unroll
            for (uint32_t x = 0; x < kXcds; ++x) {
                const float2 v = *reinterpret_cast<const float2*>(grad_x + 2 * (x * entries + e));
                a = a + v.x;
                b = b + v.y;
            }
        if (use_xh)
#pragma unroll
            for (uint32_t x = 0; x < kXcds; ++x) {
                const __half2 h = grad_xh[x * entries + e];
                a = a + __low2float(h);
                b = b + __high2float(h);
            }
        if (accumulate) {
            a = a + grad_tab[2 * e];
            b = b + grad_tab[2 * e + 1];
        }
        grad_tab[2 * e] = a;
        grad_tab[2 * e + 1] = b;
    }
}

// Sum of v over the run of equal keys that starts at this lane, in the run's
// first lane (other lanes' sums are partial).  Runs are maximal stretches of
// consecutive lanes with one key; every lane of the wave is active.
INSTA_DEV float2 run_sum(uint32_t key, float2 v, bool& head) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t prev = __shfl_up(key, 1u);
    head = lane == 0u || prev != key;
    const uint64_t heads = __ballot(head);
    const uint64_t later = lane == 63u ? 0ull : heads & (~0ull << (lane + 1u));
    const uint32_t end = later ? (uint32_t)__builtin_ctzll(later) - 1u : 63u;
#pragma unroll
    for (uint32_t off = 1; off < 64u; off <<= 1) {
        const float ox = __shfl_down(v.x, off), oy = __shfl_down(v.y, off);
        if (lane + off <= end) {
            v.x = v.x + ox;
            v.y = v.y + oy;
        }
    }
    return v;
}

// Hash-grid scatter of the leading levels whose tables fit in LDS (at most
// kLdsEntries entries each): one workgroup per (level, chunk of the batch)
// adds its samples' corner gradients into an LDS copy of the level's table
// (LDS float atomics, or packed FP16 atomics under H16, one per corner
// instead of two; on levels in run_mask a wave first sums runs of
// consecutive samples with one corner, as the per-sample scatter does), then
// stores the whole table as its chunk's partial (plain coalesced stores, no
// global atomics); grad_copies_reduce_kernel sums the partials in chunk
// order.  Same corner and weight arithmetic as the per-sample scatter.
constexpr uint32_t kLdsEntries = 16384;
constexpr uint32_t kScatterThreads = 1024;
constexpr uint32_t kScatterChunk = 8192;    // samples per workgroup
constexpr uint32_t kScatterMaxChunks = 64;  // partial tables

// The leading levels whose tables fit in LDS, as a mask, and the first entry
// after them.
uint32_t lds_scatter_levels(const GridParams& gp, int n_levels, uint32_t* end) {
    uint32_t mask = 0, e = 0;
    for (int l = 0; l < n_levels && gp.size[l] <= kLdsEntries; ++l) {
        mask |= 1u << l;
        e = gp.offset[l] + gp.size[l];
    }
    *end = e;
    return mask;
}

template <bool H16>
__global__ __launch_bounds__(kScatterThreads) void scatter_lds_kernel(GridParams gp, const uint32_t* __restrict__ n_dev,
                                                                      const uint32_t* __restrict__ c_idx,
                                                                      const float4* __restrict__ cpos,
                                                                      const float* __restrict__ denc_g, uint32_t ld,
                                                                      uint32_t run_mask, float* __restrict__ grad_p,
                                                                      uint32_t lds_end) {
    // FP32 pairs (128 KB, one workgroup per CU) or packed FP16 (64 KB, two)
    using Entry = std::conditional_t<H16, __half2, float2>;
    __shared__ Entry tab[kLdsEntries];
    const uint32_t l = blockIdx.y;  // grid.y: the LDS-scattered levels
    const uint32_t n = *n_dev;
    const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    // an empty chunk still stores its (zero) partial
    const uint32_t s0 = min(n, blockIdx.x * chunk), s1 = min(n, s0 + chunk);
    const uint32_t size = gp.size[l], off = gp.offset[l], res = gp.res[l];
    const bool dense_l = (gp.dense_mask >> l) & 1u;
    const float sc = gp.scale[l];
    for (uint32_t j = threadIdx.x; j < size; j += kScatterThreads) {
        if constexpr (H16)
            tab[j] = __floats2half2_rn(0.0f, 0.0f);
        else
            tab[j] = make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const bool runs = (run_mask >> l) & 1u;
    // every thread runs every round (run_sum needs whole waves); the next
    // round's sample is loaded before this round's corners are added
    const uint32_t rounds = (s1 - s0 + kScatterThreads - 1) / kScatterThreads;
    auto load = [&](uint32_t q, float4& p4, float& e0, float& e1) {
        const uint32_t i = s0 + q * kScatterThreads + threadIdx.x;
        if (i < s1) {
            p4 = cpos[c_idx[i]];
            e0 = denc_g[(size_t)(2 * l) * ld + i];
            e1 = denc_g[(size_t)(2 * l + 1) * ld + i];
        } else {
            p4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            e0 = e1 = 0.0f;
        }
    };
    float4 pn;
    float en0 = 0.0f, en1 = 0.0f;
    if (rounds) load(0, pn, en0, en1);
    for (uint32_t q = 0; q < rounds; ++q) {
        const bool valid = s0 + q * kScatterThreads + threadIdx.x < s1;
        const float xs[3] = {pn.x, pn.y, pn.z};
        const float e0 = en0, e1 = en1;
        if (q + 1 < rounds) load(q + 1, pn, en0, en1);
        float frac[3];
        uint32_t g[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float pp = xs[d] * sc + 0.5f;
            const float fl = floorf(pp);
            frac[d] = pp - fl;
            g[d] = (uint32_t)(int32_t)fl;
        }
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) {
            float w = 1.0f;
            uint32_t cg[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const bool up = (c >> k) & 1u;
                w = w * (up ? frac[k] : 1.0f - frac[k]);
                cg[k] = g[k] + (up ? 1u : 0u);
            }
            const uint32_t idx = grid_index(cg, dense_l, res, size);
            float2 v = make_float2(w * e0, w * e1);
            bool lead = valid;
            if (runs) {
                bool head;
                v = run_sum(valid ? idx : ~0u, v, head);
                lead = valid && head;
            }
            if (lead) {
                if constexpr (H16) {
                    unsafeAtomicAdd(&tab[idx], __floats2half2_rn(v.x, v.y));
                } else {
                    atomicAdd(&tab[idx].x, v.x);
                    atomicAdd(&tab[idx].y, v.y);
                }
            }
        }
    }
    __syncthreads();
    float* __restrict__ dst = grad_p + 2u * ((size_t)blockIdx.x * lds_end + off);
    float2* __restrict__ dst2 = reinterpret_cast<float2*>(dst);
    for (uint32_t j = threadIdx.x; j < size; j += kScatterThreads) {
        if constexpr (H16)
            dst2[j] = __half22float2(tab[j]);
        else
            dst2[j] = tab[j];
    }
}

// Backward over a block of 256 compacted samples, in two phases.
// Phase 1 (FP32 MFMA, the forward kernel's layout, 16 samples per step of a
// wave): the deltas of both MLPs through the transposed weights (held in
// LDS), the ReLU masks from the stored activations, each delta's rows stored
// for the weight gradients, and dL/dencoding into LDS.  Phase 2 (one thread
// per sample): the hash-grid scatter; on the levels of run_mask the corner
// gradients of consecutive samples that land on one table entry (coarse
// levels along one ray) are summed across the wave first and added by one
// lane.  (A split into a delta kernel and a low-register scatter kernel
// measured slower: the scatter's extra occupancy does not repay the round
// trip of dL/dencoding through memory.)
template <int L, bool H16, int X>
__global__ __launch_bounds__(kMlpThreads) void train_backward_kernel(const float* __restrict__ params, GridParams gp,
                                                                       const uint32_t* __restrict__ n_dev,
                                                                       const uint32_t* __restrict__ c_idx,
                                                                       const float4* __restrict__ cpos,
                                                                       const float* __restrict__ act,
                                                                       const float4* __restrict__ dout,
                                                                       float* __restrict__ dact, uint32_t ld,
                                                                       float* __restrict__ grad_x,
                                                                       __half2* __restrict__ grad_xh,
                                                                       uint64_t entries, uint32_t run_mask,
                                                                       float* __restrict__ denc_g,
                                                                       uint32_t lds_mask) {
    constexpr int E = 2 * L, DIN = E + X;  // encoding, density input (+ extra dims)
    using A = Act<DIN>;
    using O = Off<DIN>;
    typedef float bf4 __attribute__((ext_vector_type(4)));
    __shared__ float w[O::N];                  // both MLPs' weights
    __shared__ float s_denc[E][kMlpThreads];   // dL/dencoding of the block's samples
    const uint32_t n = *n_dev;
    const uint32_t base = blockIdx.x * kMlpThreads;
    if (base >= n) return;  // whole block past the batch
    for (uint32_t e = threadIdx.x; e < O::N; e += kMlpThreads) w[e] = params[e];
    __syncthreads();
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t m = lane & 15u, g = lane >> 4;
    for (uint32_t nb = 0; nb < 4; ++nb) {
        // Opaque copies of 0 and ld: the weight reads (LDS) and the row
        // addresses of act/dact are formed inside this loop.  Loop-invariant,
        // they were otherwise hoisted out of it into ~360 registers (1 wave
        // per SIMD); recomputed per block of 16 samples, 248 (2 waves).
        uint32_t z = 0, ldz = ld;
        asm volatile("" : "+s"(z), "+s"(ldz));
        const float* W0 = w + O::W0 + z;
        const float* W1 = w + O::W1 + z;
        const float* W2 = w + O::W2 + z;
        const float* W3 = w + O::W3 + z;
        const float* W4 = w + O::W4 + z;
        const uint32_t s = wave * 64 + nb * 16 + m;  // column (sample) of this lane
        const uint32_t i = base + s;
        const bool live = i < n;
        const float4 d4 = live ? dout[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        // colour output layer: dH3 = relu'(H3) (W4^T dRGB), K = 3 rgb rows (+ one zero)
        const float b_rgb = g == 0 ? d4.x : (g == 1 ? d4.y : (g == 2 ? d4.z : 0.0f));
        bf4 dh3[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(g < 3 ? W4[g * kW + 16 * t + m] : 0.0f, b_rgb, acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t row = 16 * t + 4 * g + q;
                dh3[t][q] = live && act[(size_t)(A::H3 + row) * ldz + i] > 0.0f ? acc[q] : 0.0f;
                if (live) dact[(size_t)(Dact::DH3 + row) * ldz + i] = dh3[t][q];
            }
        }
        // colour hidden layer: dH2 = relu'(H2) (W3^T dH3)
        bf4 dh2[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W3[(16 * s2 + 4 * g + b) * kW + 16 * t + m], dh3[s2][b],
                                                               acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t row = 16 * t + 4 * g + q;
                dh2[t][q] = live && act[(size_t)(A::H2 + row) * ldz + i] > 0.0f ? acc[q] : 0.0f;
                if (live) dact[(size_t)(Dact::DH2 + row) * ldz + i] = dh2[t][q];
            }
        }
        // colour input restricted to the density outputs: dD = (W2^T dH2)[0..15], + dL/d(density logit) on row 0
        bf4 dd = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                dd = __builtin_amdgcn_mfma_f32_16x16x4f32(W2[(16 * t + 4 * g + b) * kCin + m], dh2[t][b], dd, 0, 0, 0);
        if (g == 0) dd[0] = dd[0] + d4.w;  // density logit = density-MLP output 0
        if (live)
#pragma unroll
            for (int q = 0; q < 4; ++q) dact[(size_t)(Dact::DDENS + 4 * g + q) * ldz + i] = dd[q];
        // density hidden layer: dH1 = relu'(H1) (W1^T dD)
        bf4 dh1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int b = 0; b < 4; ++b)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W1[(4 * g + b) * kW + 16 * t + m], dd[b], acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t row = 16 * t + 4 * g + q;
                dh1[t][q] = live && act[(size_t)(A::H1 + row) * ldz + i] > 0.0f ? acc[q] : 0.0f;
                if (live) dact[(size_t)(Dact::DH1 + row) * ldz + i] = dh1[t][q];
            }
        }
        // dL/dencoding = (W0^T dH1)[0..E), the extra dims' rows not needed
#pragma unroll
        for (int te = 0; te < E / 16; ++te) {
            bf4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W0[(16 * s2 + 4 * g + b) * DIN + 16 * te + m],
                                                               dh1[s2][b], acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) s_denc[16 * te + 4 * g + q][s] = acc[q];
        }
    }
    __syncthreads();
    const uint32_t i = base + tid;
    if ((i & ~63u) >= n) return;  // whole wave past the batch (no barrier follows)
    const bool valid = i < n;
    const uint32_t ic = valid ? i : n - 1u;  // tail lanes recompute the last sample and store nothing
    if (valid) {
        const float4 d4 = dout[i];
        dact[(size_t)(Dact::DRGB + 0) * ld + i] = d4.x;
        dact[(size_t)(Dact::DRGB + 1) * ld + i] = d4.y;
        dact[(size_t)(Dact::DRGB + 2) * ld + i] = d4.z;
    }
    // hash-grid scatter (the corner/weight arithmetic of encode_level) into
    // this XCD's copy of the table gradients: a copy's lines stay in one
    // XCD's L2, instead of moving between the L2s of the eight XCDs that
    // would otherwise add to the same entries
    const uint32_t xcd = xcd_id();
    float* __restrict__ grad_tab = grad_x + (size_t)xcd * 2u * entries;
    __half2* __restrict__ grad_tab_h = grad_xh + (size_t)xcd * entries;
    const float4 p4 = cpos[c_idx[ic]];
    const float xs[3] = {p4.x, p4.y, p4.z};
#pragma unroll 1
    for (int l = 0; l < L; ++l) {
        if ((lds_mask >> l) & 1u) {
            // scattered by scatter_lds_kernel, which reads dL/dencoding from memory
            if (valid) {
                denc_g[(size_t)(2 * l) * ld + i] = s_denc[2 * l][threadIdx.x];
                denc_g[(size_t)(2 * l + 1) * ld + i] = s_denc[2 * l + 1][threadIdx.x];
            }
            continue;
        }
        const float sc = gp.scale[l];
        float frac[3];
        uint32_t g[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float pp = xs[d] * sc + 0.5f;
            const float fl = floorf(pp);
            frac[d] = pp - fl;
            g[d] = (uint32_t)(int32_t)fl;
        }
        const bool dense_l = (gp.dense_mask >> l) & 1u;
        const bool runs = (run_mask >> l) & 1u;
        const uint32_t res = gp.res[l], size = gp.size[l], off = gp.offset[l];
        const float e0 = valid ? s_denc[2 * l][threadIdx.x] : 0.0f;
        const float e1 = valid ? s_denc[2 * l + 1][threadIdx.x] : 0.0f;
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) {
            float w = 1.0f;
            uint32_t cg[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const bool up = (c >> k) & 1u;
                w = w * (up ? frac[k] : 1.0f - frac[k]);
                cg[k] = g[k] + (up ? 1u : 0u);
            }
            const uint32_t idx = off + grid_index(cg, dense_l, res, size);
            float2 v = make_float2(w * e0, w * e1);
            bool lead = valid;
            if (runs) {
                bool head;
                v = run_sum(valid ? idx : ~0u, v, head);
                lead = valid && head;
            }
            if (lead) {
                if constexpr (H16) {
                    unsafeAtomicAdd(grad_tab_h + idx, __floats2half2_rn(v.x, v.y));
                } else {
                    atomicAdd(grad_tab + 2 * (size_t)idx, v.x);
                    atomicAdd(grad_tab + 2 * (size_t)idx + 1, v.y);
                }
            }
        }
    }
}


struct WProduct {
    uint32_t d_row, x_row, O, I, out;
};
struct WArgs {
    WProduct p[5];
};

// partial[c][out + o*I + i] = sum over samples n of chunk c of
// dact[d_row + o][n] * act[x_row + i][n], on v_mfma_f32_16x16x4f32: the
// samples are the K dimension; kS-sample slices of both operands are staged
// through LDS (each thread loads 4 consecutive samples of kPass rows of
// each operand as float4, coalesced along the samples), wave w owns input
// tile i = 16w..16w+15 and the four output tiles o = 0..63 of it, one MFMA
// per tile and 4 samples (A = 16 outputs x 4 samples, B = 4 samples x 16
// inputs).
typedef float wg_f4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void wgrad_kernel(WArgs wa, const uint32_t* __restrict__ n_dev,
                                                    const float* __restrict__ act, const float* __restrict__ dact,
                                                    uint32_t ld, float* __restrict__ part, uint32_t n_mlp) {
    constexpr uint32_t kS = 32;            // samples per LDS slice
    constexpr uint32_t kP = kS + 4;        // LDS row pitch: fragment reads (16 rows x 4 samples) spread over the banks
    constexpr uint32_t kRowsPass = 1024u / kS;  // rows one pass of 256 threads covers (kS / 4 threads per row)
    constexpr uint32_t kPass = 64u / kRowsPass;
    __shared__ float Ds[64 * kP];
    __shared__ float Xs[64 * kP];
    const WProduct pr = wa.p[blockIdx.y];
    const uint32_t n = *n_dev;
    const uint32_t s0 = blockIdx.x * kChunk;
    if (s0 >= n) return;
    const uint32_t s1 = min(s0 + kChunk, n);
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t m = lane & 15u, g = lane >> 4;   // fragment row / k of this lane
    const uint32_t n_ot = (pr.O + 15u) / 16u;       // output tiles in use
    const bool active = 16u * wave < pr.I;          // this wave's input tile exists
    const bool vec = (ld & 3u) == 0u;               // rows 16-byte aligned
    const uint32_t c4 = (t % (kS / 4u)) * 4u;       // this thread's first sample in a slice
    const uint32_t r0 = t / (kS / 4u);              // its rows r0 + kRowsPass p
    wg_f4 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = wg_f4{0.0f, 0.0f, 0.0f, 0.0f};
    const float* dp[kPass];
    const float* xp[kPass];
#pragma unroll
    for (uint32_t q = 0; q < kPass; ++q) {
        const uint32_t row = r0 + kRowsPass * q;
        dp[q] = row < pr.O ? dact + (size_t)(pr.d_row + row) * ld : nullptr;
        xp[q] = row < pr.I ? act + (size_t)(pr.x_row + row) * ld : nullptr;
    }
    auto ld4 = [&](const float* p, uint32_t smp) -> float4 {
        if (!p) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (vec && smp + 4u <= s1) return *reinterpret_cast<const float4*>(p + smp);
        return make_float4(smp < s1 ? p[smp] : 0.0f, smp + 1u < s1 ? p[smp + 1u] : 0.0f,
                           smp + 2u < s1 ? p[smp + 2u] : 0.0f, smp + 3u < s1 ? p[smp + 3u] : 0.0f);
    };
    // next slice in registers while the current one is multiplied
    float4 dn[kPass], xn[kPass];
    auto fetch = [&](uint32_t s) {
#pragma unroll
        for (uint32_t q = 0; q < kPass; ++q) {
            dn[q] = ld4(dp[q], s + c4);
            xn[q] = ld4(xp[q], s + c4);
        }
    };
    fetch(s0);
    for (uint32_t s = s0; s < s1; s += kS) {
#pragma unroll
        for (uint32_t q = 0; q < kPass; ++q) {
            const uint32_t row = r0 + kRowsPass * q;
            *reinterpret_cast<float4*>(&Ds[row * kP + c4]) = dn[q];
            *reinterpret_cast<float4*>(&Xs[row * kP + c4]) = xn[q];
        }
        __syncthreads();
        if (s + kS < s1) fetch(s + kS);
        // one output tile (O <= 16) or all four (rows past O are zero):
        // straight-line MFMAs, the accumulators stay in place
        auto mma = [&](auto tiles) {
#pragma unroll
            for (uint32_t k0 = 0; k0 < kS; k0 += 4u) {
                const float b = Xs[(16u * wave + m) * kP + k0 + g];
#pragma unroll
                for (uint32_t ot = 0; ot < decltype(tiles)::value; ++ot)
                    acc[ot] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ds[(16u * ot + m) * kP + k0 + g], b, acc[ot],
                                                                    0, 0, 0);
            }
        };
        if (active) {
            if (n_ot <= 1u)
                mma(std::integral_constant<uint32_t, 1>{});
            else
                mma(std::integral_constant<uint32_t, 4>{});
        }
        __syncthreads();
    }
    if (!active) return;
    // lane holds C[o = 16 ot + 4 g + j][i = 16 wave + m]
    const uint32_t ii = 16u * wave + m;
#pragma unroll
    for (uint32_t ot = 0; ot < 4u; ++ot)
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j) {
            const uint32_t o = 16u * ot + 4u * g + j;
            if (o < pr.O && ii < pr.I) part[(size_t)blockIdx.x * n_mlp + pr.out + o * pr.I + ii] = acc[ot][j];
        }
}

__global__ void wgrad_reduce_kernel(const float* __restrict__ part, const uint32_t* __restrict__ n_dev,
                                    uint32_t n_mlp, float* __restrict__ grad) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_mlp) return;
    const uint32_t nc = (*n_dev + kChunk - 1) / kChunk;
    // eight interleaved running sums (chunk c into sum c mod 8), added in a
    // fixed order: deterministic, with eight independent loads in flight
    float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t c = 0;
    for (; c + 8 <= nc; c += 8)
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r) s[r] += part[(size_t)(c + r) * n_mlp + p];
#pragma unroll
    for (uint32_t r = 0; r < 8; ++r)
        if (c + r < nc) s[r] += part[(size_t)(c + r) * n_mlp + p];
    grad[p] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// Unscale, optionally step Adam.  grad[] receives dL/dparam (unscaled).
__global__ void adam_kernel(float* __restrict__ params, float* __restrict__ grad, const __half2* __restrict__ grad_h,
                            float* __restrict__ m, float* __restrict__ v, uint64_t n_params, uint64_t n_mlp,
                            float inv_scale, float lr, float b1, float b2, float eps, float bc1, float bc2,
                            float l2_reg, uint32_t apply) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n_params;
         p += (uint64_t)gridDim.x * blockDim.x) {
        float g;
        if (grad_h && p >= n_mlp) {
            const uint64_t q = p - n_mlp;
            const __half2 h = grad_h[q >> 1];
            g = (q & 1u) ? __high2float(h) : __low2float(h);
        } else {
            g = grad[p];
        }
        g = g * inv_scale;
        grad[p] = g;
        if (apply) {
            g = g + l2_reg * params[p];
            const float mm = b1 * m[p] + (1.0f - b1) * g;
            const float vv = b2 * v[p] + (1.0f - b2) * (g * g);
            m[p] = mm;
            v[p] = vv;
            params[p] = params[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
        }
    }
}

// Ema around the optimizer: a = decay * a + (1 - decay) * w (a starts at 0
// with the first update), and the debiased average a * debias, debias =
// 1 / (1 - decay^t), for inference
__global__ void ema_kernel(const float* __restrict__ w, float* __restrict__ a, float* __restrict__ infer,
                           uint64_t n, float decay, float debias, uint32_t first) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x) {
        const float prev = first ? 0.0f : a[p];
        const float v = decay * prev + (1.0f - decay) * w[p];
        a[p] = v;
        infer[p] = v * debias;
    }
}

template <class T>
bool train_alloc(T*& p, size_t bytes) {
    return hipMalloc(reinterpret_cast<void**>(&p), bytes) == hipSuccess;
}

void free_workspace_train(TrainState* ts) {
    dev_free(ts->ray_d); dev_free(ts->ray_o); dev_free(ts->ray_frame); dev_free(ts->ray_img); dev_free(ts->ray_pix);
    dev_free(ts->ray_src); dev_free(ts->ray_base); dev_free(ts->ray_cnt);
    dev_free(ts->s_cpos); dev_free(ts->s_cdir); dev_free(ts->s_out); dev_free(ts->s_hit);
    dev_free(ts->c_idx); dev_free(ts->c_dout); dev_free(ts->act); dev_free(ts->dact); dev_free(ts->part);
    dev_free(ts->t_buf);
    dev_free(ts->denc);
    dev_free(ts->ray_T); dev_free(ts->ray_alive); dev_free(ts->ray_gbase);
    dev_free(ts->r_pos); dev_free(ts->r_cpos); dev_free(ts->r_cdir); dev_free(ts->r_out);
    ts->ray_cap = ts->sample_cap = ts->compact_cap = 0;
}

// rays, marched samples and the compacted batch grow independently to the
// largest request seen
// grow the training workspace to hold n_rays rays, n_samples marched samples
// and a compacted batch of n_compact (insta_reserve only)
insta_status grow_train_workspace(insta_ctx* ctx, TrainState* ts, uint32_t n_rays, uint32_t n_samples,
                                  uint32_t n_compact) {
    if (n_rays <= ts->ray_cap && n_samples <= ts->sample_cap && n_compact <= ts->compact_cap) return INSTA_OK;
    const uint32_t rc = std::max(n_rays, ts->ray_cap), sc = std::max(n_samples, ts->sample_cap);
    const uint32_t cc = std::max(n_compact, ts->compact_cap);
    free_workspace_train(ts);
    const int DIN = (int)(ctx->cfg.n_levels * 2 + ctx->cfg.n_extra_dims);
    const size_t act_rows = (size_t)DIN + 224, n_mlp = ctx->layout.n_mlp_params;
    bool ok = true;
    ok = ok && train_alloc(ts->ray_d, rc * sizeof(float4)) && train_alloc(ts->ray_o, rc * sizeof(float4)) &&
         train_alloc(ts->ray_frame, rc * sizeof(int32_t));
    ok = ok && train_alloc(ts->ray_img, rc * 4) && train_alloc(ts->ray_pix, rc * 4) && train_alloc(ts->ray_src, rc * 4);
    ok = ok && train_alloc(ts->ray_base, rc * 4) && train_alloc(ts->ray_cnt, rc * 4);
    ok = ok && train_alloc(ts->s_cpos, sc * sizeof(float4));
    ok = ok && train_alloc(ts->s_cdir, sc * sizeof(float4)) && train_alloc(ts->s_out, sc * sizeof(float4));
    ok = ok && train_alloc(ts->s_hit, sc * sizeof(int32_t));
    ok = ok && train_alloc(ts->ray_T, rc * sizeof(float)) && train_alloc(ts->ray_alive, rc * 4) &&
         train_alloc(ts->ray_gbase, rc * 4);
    ok = ok && train_alloc(ts->r_pos, sc * sizeof(float4)) && train_alloc(ts->r_cpos, sc * sizeof(float4));
    ok = ok && train_alloc(ts->r_cdir, sc * sizeof(float4)) && train_alloc(ts->r_out, sc * sizeof(float4));
    ok = ok && train_alloc(ts->c_idx, cc * sizeof(uint32_t)) && train_alloc(ts->c_dout, cc * sizeof(float4));
    ok = ok && train_alloc(ts->act, act_rows * cc * sizeof(float));
    ok = ok && train_alloc(ts->dact, (size_t)Dact::N * cc * sizeof(float));
    ok = ok && train_alloc(ts->part, ((cc + kChunk - 1) / kChunk) * n_mlp * sizeof(float));
    ok = ok && train_alloc(ts->t_buf, (size_t)rc * ctx->cfg.max_steps * sizeof(float));
    ok = ok && train_alloc(ts->denc, (size_t)2 * INSTA_MAX_LEVELS * cc * sizeof(float));
    if (!ok) {
        free_workspace_train(ts);
        set_error("training workspace allocation failed");
        return INSTA_ERR_OOM;
    }
    ts->ray_cap = rc;
    ts->sample_cap = sc;
    ts->compact_cap = cc;
    return INSTA_OK;
}

// a step or plugin call must fit the reserved workspace: nothing is allocated per call
insta_status check_train_workspace(const TrainState* ts, uint32_t n_rays, uint32_t n_samples, uint32_t n_compact,
                                   const char* who) {
    if (n_rays <= ts->ray_cap && n_samples <= ts->sample_cap && n_compact <= ts->compact_cap) return INSTA_OK;
    set_error(std::string(who) + ": needs " + std::to_string(n_rays) + " rays, " + std::to_string(n_samples) +
              " samples, a batch of " + std::to_string(n_compact) + "; reserved " + std::to_string(ts->ray_cap) +
              " / " + std::to_string(ts->sample_cap) + " / " + std::to_string(ts->compact_cap) +
              ": reserve them with insta_reserve (train_rays, train_samples, train_batch)");
    return INSTA_ERR_STATE;
}

template <int L, int X>
insta_status launch_train_forward(insta_ctx* ctx, TrainState* ts, hipStream_t stream) {
    const uint32_t cap = ts->compact_cap;
    const uint32_t* n_dev = ts->counters + kTCompactN;
    const float* table = ctx->d_params + ctx->layout.n_mlp_params;
    const GridParams gp = grid_params(ctx);
    const dim3 sgrid((cap + kFwdTile - 1) / kFwdTile);
    hipLaunchKernelGGL((train_forward_kernel<L, X>), sgrid, dim3(kFwdTile), 0, stream, ctx->d_params, table, gp,
                       ctx->d_frame_extra, ctx->cfg.max_frames, n_dev, ts->c_idx, ts->s_cpos, ts->s_cdir,
                       static_cast<const float*>(nullptr), ts->act, cap, static_cast<float4*>(nullptr));
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

template <int L, int X>
insta_status launch_train_backward(insta_ctx* ctx, TrainState* ts, const insta_train_config& tc, hipStream_t stream) {
    constexpr int DIN = 2 * L + X;
    using A = Act<DIN>;
    using O = Off<DIN>;
    const uint32_t cap = ts->compact_cap;
    const uint32_t* n_dev = ts->counters + kTCompactN;
    const GridParams gp = grid_params(ctx);
    const uint32_t n_mlp = (uint32_t)ctx->layout.n_mlp_params;
    const dim3 sgrid((cap + kMlpThreads - 1) / kMlpThreads);
    float* grad_tab = ts->grad + n_mlp;
    const uint64_t entries = ctx->layout.n_grid_params / 2;
    // levels whose tables fit in LDS are scattered by scatter_lds_kernel,
    // the others per sample (FP32, or packed FP16 copies)
    const uint32_t lds_mask = ts->lds_levels, lds_end = ts->lds_end;
    const bool per_sample = lds_mask != (1u << L) - 1u;
    const bool use_xh = tc.grid_grad_fp16 && per_sample;
    const bool use_x = !tc.grid_grad_fp16 && per_sample;
    // clear this backward's copies (fill kernels: capturable)
    if (use_xh)
        hipLaunchKernelGGL(fill_zero_kernel, dim3(2048), dim3(256), 0, stream, reinterpret_cast<uint32_t*>(ts->grad_xh),
                           kXcds * entries);
    if (use_x)
        hipLaunchKernelGGL(fill_zero_kernel, dim3(2048), dim3(256), 0, stream, reinterpret_cast<uint32_t*>(ts->grad_x),
                           kXcds * 2 * entries);
    // levels whose cells span at least 1 / kRunMaxStepCells marching steps
    uint32_t run_mask = 0;
    for (int l = 0; l < L; ++l)
        if (ctx->cfg.dt * gp.scale[l] <= kRunMaxStepCells) run_mask |= 1u << l;
    if (tc.grid_grad_fp16)
        hipLaunchKernelGGL((train_backward_kernel<L, true, X>), sgrid, dim3(kMlpThreads), 0, stream, ctx->d_params, gp,
                           n_dev, ts->c_idx, ts->s_cpos, ts->act, ts->c_dout, ts->dact, cap, ts->grad_x, ts->grad_xh,
                           entries, run_mask, ts->denc, lds_mask);
    else
        hipLaunchKernelGGL((train_backward_kernel<L, false, X>), sgrid, dim3(kMlpThreads), 0, stream, ctx->d_params, gp,
                           n_dev, ts->c_idx, ts->s_cpos, ts->act, ts->c_dout, ts->dact, cap, ts->grad_x, ts->grad_xh,
                           entries, run_mask, ts->denc, lds_mask);
    INSTA_HIP(hipGetLastError());
    const uint32_t chunks =
        std::min<uint32_t>(std::max<uint32_t>((cap + kScatterChunk - 1) / kScatterChunk, 1u), kScatterMaxChunks);
    if (lds_mask) {
        const dim3 grid(chunks, __builtin_popcount(lds_mask));
        if (tc.grid_grad_fp16)
            hipLaunchKernelGGL(scatter_lds_kernel<true>, grid, dim3(kScatterThreads), 0, stream, gp, n_dev, ts->c_idx,
                               ts->s_cpos, ts->denc, cap, run_mask, ts->grad_p, lds_end);
        else
            hipLaunchKernelGGL(scatter_lds_kernel<false>, grid, dim3(kScatterThreads), 0, stream, gp, n_dev, ts->c_idx,
                               ts->s_cpos, ts->denc, cap, run_mask, ts->grad_p, lds_end);
        INSTA_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(grad_copies_reduce_kernel, dim3(2048), dim3(256), 0, stream, ts->grad_x, ts->grad_xh, entries,
                       use_x ? 1u : 0u, use_xh ? 1u : 0u, ts->grad_p, lds_end, chunks, grad_tab, 0u);
    INSTA_HIP(hipGetLastError());
    WArgs wa;
    wa.p[0] = WProduct{Dact::DH1, A::ENC, kW, DIN, O::W0};
    wa.p[1] = WProduct{Dact::DDENS, A::H1, kDOut, kW, O::W1};
    wa.p[2] = WProduct{Dact::DH2, A::DENS, kW, kCin, O::W2};
    wa.p[3] = WProduct{Dact::DH3, A::H2, kW, kW, O::W3};
    wa.p[4] = WProduct{Dact::DRGB, A::H3, 3, kW, O::W4};
    const uint32_t n_chunks = (cap + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(wgrad_kernel, dim3(n_chunks, 5), dim3(256), 0, stream, wa, n_dev, ts->act, ts->dact, cap,
                       ts->part, n_mlp);
    INSTA_HIP(hipGetLastError());
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((n_mlp + 255) / 256), dim3(256), 0, stream, ts->part, n_dev, n_mlp,
                       ts->grad);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

// ---- NerfNetwork plugin (nerf_network.h:238-383): forward with stored
// activations and backward on caller-given samples, through the training
// step's kernels; the batch is the caller's samples in order (c_idx = i).
__global__ void net_prepare_kernel(const float* __restrict__ coords, uint32_t n, uint32_t stride, uint32_t X,
                                   float4* __restrict__ cpos, float4* __restrict__ cdir, uint32_t* __restrict__ c_idx,
                                   float* __restrict__ x_extra, uint32_t* __restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) counters[kTCompactN] = n;
    if (i >= n) return;
    const float* c = coords + (size_t)i * stride;
    cpos[i] = make_float4(c[0], c[1], c[2], 0.0f);
    cdir[i] = make_float4(c[8], c[9], c[10], __uint_as_float(kExtraOnes));
    c_idx[i] = i;
    // the density input's extra dims are the coordinate's last X floats
    // (forward_impl's extract_expressions, nerf_network.h:253-262)
    for (uint32_t e = 0; e < X; ++e) x_extra[(size_t)i * X + e] = c[stride - X + e];
}

// param_gradients_mode Overwrite / Accumulate
__global__ void net_grad_out_kernel(const float* __restrict__ g, float* __restrict__ out, uint64_t n,
                                    uint32_t accumulate) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x)
        out[p] = accumulate ? out[p] + g[p] : g[p];
}

// Input gradients of the backward (dL_dinput rows 0..2 and the direction rows
// 8..10; the reference backpropagates through the position and direction
// encodings only, nerf_network.h:318-380): dL/dencoding = W0^T dH1 and
// dL/dSH = W2[:, 16:32]^T dH2 from the stored deltas; the position through
// the derivative of each level's trilinear weights (scale_l per axis) times
// the corner features, the direction (input in [0,1], d = 2 x - 1) through
// the derivatives of the 16 SH polynomials.
INSTA_DEV void sh16_grad(float x, float y, float z, const float* g, float out[3]) {
    const float C1 = 0.48860251190291987f, C2 = 1.0925484305920792f, C3 = 0.94617469575755997f;
    const float C4 = 0.54627421529603959f, C5 = 0.59004358992664352f, C6 = 2.8906114426405538f;
    const float C7 = 0.45704579946446572f, C8 = 0.3731763325901154f, C9 = 1.4453057213202769f;
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
    dy += g[1] * -C1;
    dz += g[2] * C1;
    dx += g[3] * -C1;
    dx += g[4] * (C2 * y); dy += g[4] * (C2 * x);
    dy += g[5] * (-C2 * z); dz += g[5] * (-C2 * y);
    dz += g[6] * (2.0f * C3 * z);
    dx += g[7] * (-C2 * z); dz += g[7] * (-C2 * x);
    dx += g[8] * (2.0f * C4 * x); dy += g[8] * (-2.0f * C4 * y);
    dx += g[9] * (-6.0f * C5 * x * y); dy += g[9] * (C5 * (3.0f * y * y - 3.0f * x * x));
    dx += g[10] * (C6 * y * z); dy += g[10] * (C6 * x * z); dz += g[10] * (C6 * x * y);
    dy += g[11] * (C7 * (1.0f - 5.0f * z * z)); dz += g[11] * (-10.0f * C7 * y * z);
    dz += g[12] * (C8 * (15.0f * z * z - 3.0f));
    dx += g[13] * (C7 * (1.0f - 5.0f * z * z)); dz += g[13] * (-10.0f * C7 * x * z);
    dx += g[14] * (2.0f * C9 * x * z); dy += g[14] * (-2.0f * C9 * y * z); dz += g[14] * (C9 * (x * x - y * y));
    dx += g[15] * (C5 * (3.0f * y * y - 3.0f * x * x)); dy += g[15] * (6.0f * C5 * x * y);
    out[0] = dx;
    out[1] = dy;
    out[2] = dz;
}

template <int L, int X>
__global__ __launch_bounds__(128) void net_input_grad_kernel(const float* __restrict__ params,
                                                             const float* __restrict__ table, GridParams gp,
                                                             uint32_t n, const float4* __restrict__ cpos,
                                                             const float4* __restrict__ cdir,
                                                             const float* __restrict__ dact, uint32_t ld,
                                                             float* __restrict__ dL_dinput, uint32_t stride) {
    constexpr int E = 2 * L, DIN = E + X;
    using O = Off<DIN>;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float denc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) denc[e] = 0.0f;
    float dsh[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) dsh[j] = 0.0f;
    for (int k = 0; k < kW; ++k) {
        const float d1 = dact[(size_t)(Dact::DH1 + k) * ld + i];
        const float d2 = dact[(size_t)(Dact::DH2 + k) * ld + i];
#pragma unroll
        for (int e = 0; e < E; ++e) denc[e] += params[O::W0 + k * DIN + e] * d1;
#pragma unroll
        for (int j = 0; j < 16; ++j) dsh[j] += params[O::W2 + k * kCin + 16 + j] * d2;
    }
    const float4 p4 = cpos[i];
    const float xs[3] = {p4.x, p4.y, p4.z};
    float gx[3] = {0.0f, 0.0f, 0.0f};
    for (int l = 0; l < L; ++l) {
        const float sc = gp.scale[l];
        float frac[3];
        uint32_t g0[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float pp = xs[d] * sc + 0.5f;
            const float fl = floorf(pp);
            frac[d] = pp - fl;
            g0[d] = (uint32_t)(int32_t)fl;
        }
        const bool dense_l = (gp.dense_mask >> l) & 1u;
#pragma unroll
        for (uint32_t c = 0; c < 8; ++c) {
            uint32_t cg[3];
            float wk[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const bool up = (c >> k) & 1u;
                wk[k] = up ? frac[k] : 1.0f - frac[k];
                cg[k] = g0[k] + (up ? 1u : 0u);
            }
            const uint32_t idx = gp.offset[l] + grid_index(cg, dense_l, gp.res[l], gp.size[l]);
            const float s = denc[2 * l] * table[2 * (size_t)idx] + denc[2 * l + 1] * table[2 * (size_t)idx + 1];
            gx[0] += s * (((c & 1u) ? sc : -sc) * wk[1] * wk[2]);
            gx[1] += s * (((c & 2u) ? sc : -sc) * wk[0] * wk[2]);
            gx[2] += s * (((c & 4u) ? sc : -sc) * wk[0] * wk[1]);
        }
    }
    const float4 d4 = cdir[i];
    float gd[3];
    sh16_grad(d4.x * 2.0f - 1.0f, d4.y * 2.0f - 1.0f, d4.z * 2.0f - 1.0f, dsh, gd);
    float* o = dL_dinput + (size_t)i * stride;
    o[0] = gx[0];
    o[1] = gx[1];
    o[2] = gx[2];
    o[8] = 2.0f * gd[0];
    o[9] = 2.0f * gd[1];
    o[10] = 2.0f * gd[2];
}

template <int L, int X>
insta_status launch_net_forward(insta_ctx* ctx, TrainState* ts, float4* out, hipStream_t stream) {
    const uint32_t cap = ts->compact_cap;
    const float* table = ctx->d_params + ctx->layout.n_mlp_params;
    hipLaunchKernelGGL((train_forward_kernel<L, X>), dim3((cap + kFwdTile - 1) / kFwdTile), dim3(kFwdTile), 0, stream,
                       ctx->d_params, table, grid_params(ctx), ctx->d_frame_extra, ctx->cfg.max_frames,
                       ts->counters + kTCompactN, ts->c_idx, ts->s_cpos, ts->s_cdir,
                       X ? static_cast<const float*>(ts->net_extra) : nullptr, ts->act, cap, out);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

template <int L, int X>
insta_status launch_net_input_grad(insta_ctx* ctx, TrainState* ts, uint32_t n, float* dL_dinput, uint32_t stride,
                                   hipStream_t stream) {
    const float* table = ctx->d_params + ctx->layout.n_mlp_params;
    hipLaunchKernelGGL((net_input_grad_kernel<L, X>), dim3((n + 127) / 128), dim3(128), 0, stream, ctx->d_params,
                       table, grid_params(ctx), n, ts->s_cpos, ts->s_cdir, ts->dact, ts->compact_cap, dL_dinput,
                       stride);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

}  // namespace

bool train_supported(const insta_config& c) {
    return c.n_neurons == 64 && c.density_out == 16 && c.sh_degree == 4 && (c.n_extra_dims == 0 || c.n_extra_dims == 16) &&
           c.density_hidden_layers == 1 && c.rgb_hidden_layers == 2 && c.n_features_per_level == 2 &&
           (c.n_levels == 8 || c.n_levels == 16);
}

void free_train(insta_ctx* ctx) {
    TrainState* ts = ctx->train;
    if (!ts) return;
    free_workspace_train(ts);
    dev_free(ts->images); dev_free(ts->meta); dev_free(ts->seg); dev_free(ts->depth);
    dev_free(ts->grad); dev_free(ts->grad_x); dev_free(ts->grad_xh); dev_free(ts->grad_p);
    dev_free(ts->adam_m); dev_free(ts->adam_v); dev_free(ts->counters); dev_free(ts->net_extra);
    if (ctx->net_src == ts->ema_infer) ctx->net_src = nullptr;
    dev_free(ts->ema); dev_free(ts->ema_infer);
    if (ts->h_counters) (void)hipHostFree(ts->h_counters);
    delete ts;
    ctx->train = nullptr;
}

// gradients, optimizer moments and counters: created once per context
static insta_status ensure_train_state(insta_ctx* ctx, hipStream_t stream) {
    if (!ctx->train) {
        TrainState* ts = new (std::nothrow) TrainState();
        if (!ts) {
            set_error("host allocation failed");
            return INSTA_ERR_OOM;
        }
        ctx->train = ts;
        const uint64_t np = ctx->layout.n_params;
        ts->lds_levels = lds_scatter_levels(grid_params(ctx), (int)ctx->cfg.n_levels, &ts->lds_end);
        bool ok = train_alloc(ts->grad, np * sizeof(float)) && train_alloc(ts->adam_m, np * sizeof(float)) &&
                  train_alloc(ts->adam_v, np * sizeof(float)) && train_alloc(ts->ema, np * sizeof(float)) &&
                  train_alloc(ts->ema_infer, np * sizeof(float)) &&
                  train_alloc(ts->grad_x, kXcds * ctx->layout.n_grid_params * sizeof(float)) &&
                  train_alloc(ts->grad_xh, kXcds * (ctx->layout.n_grid_params / 2) * sizeof(__half2)) &&
                  (ts->lds_end == 0 ||
                   train_alloc(ts->grad_p, (size_t)kScatterMaxChunks * 2 * ts->lds_end * sizeof(float))) &&
                  train_alloc(ts->counters, kTCounters * sizeof(uint32_t)) &&
                  hipHostMalloc((void**)&ts->h_counters, kTCounters * sizeof(uint32_t), hipHostMallocDefault) ==
                      hipSuccess;
        if (!ok) {
            free_train(ctx);
            set_error("training state allocation failed");
            return INSTA_ERR_OOM;
        }
        INSTA_HIP(hipMemsetAsync(ts->adam_m, 0, np * sizeof(float), stream));
        INSTA_HIP(hipMemsetAsync(ts->adam_v, 0, np * sizeof(float), stream));
    }
    return INSTA_OK;
}

insta_status set_training_set(insta_ctx* ctx, uint32_t n_images, uint32_t width, uint32_t height, const float* rgba,
                              const insta_camera* cams, const int32_t* frames, hipStream_t stream) {
    insta_status st = ensure_train_state(ctx, stream);
    if (st != INSTA_OK) return st;
    TrainState* ts = ctx->train;
    dev_free(ts->images);
    dev_free(ts->meta);
    dev_free(ts->seg);    // a new set starts without supervision
    dev_free(ts->depth);
    const size_t px = (size_t)n_images * width * height;
    if (!train_alloc(ts->images, px * sizeof(float4)) || !train_alloc(ts->meta, n_images * sizeof(TrainImage))) {
        set_error("training image allocation failed");
        return INSTA_ERR_OOM;
    }
    std::vector<TrainImage> meta(n_images);
    for (uint32_t i = 0; i < n_images; ++i) {
        TrainImage& m = meta[i];
        for (int k = 0; k < 12; ++k) m.M[k] = cams[i].matrix[k];
        m.fx = cams[i].focal_px[0];
        m.fy = cams[i].focal_px[1];
        m.cx = cams[i].screen_center[0];
        m.cy = cams[i].screen_center[1];
        m.frame = frames[i];
        m.pad[0] = m.pad[1] = m.pad[2] = 0;
    }
    INSTA_HIP(hipMemcpyAsync(ts->images, rgba, px * sizeof(float4), hipMemcpyDeviceToDevice, stream));
    INSTA_HIP(hipMemcpyAsync(ts->meta, meta.data(), n_images * sizeof(TrainImage), hipMemcpyHostToDevice, stream));
    INSTA_HIP(hipStreamSynchronize(stream));  // `meta` is pageable host memory
    ts->frames.assign(frames, frames + n_images);
    ts->n_images = n_images;
    ts->width = width;
    ts->height = height;
    ts->rays_total = 0;
    return INSTA_OK;
}

insta_status set_training_supervision(insta_ctx* ctx, const uint8_t* seg, const float* depth, hipStream_t stream) {
    TrainState* ts = ctx->train;
    const size_t px = (size_t)ts->n_images * ts->width * ts->height;
    dev_free(ts->seg);
    dev_free(ts->depth);
    if ((seg && !train_alloc(ts->seg, px)) || (depth && !train_alloc(ts->depth, px * sizeof(float)))) {
        dev_free(ts->seg);
        dev_free(ts->depth);
        set_error("training supervision allocation failed");
        return INSTA_ERR_OOM;
    }
    if (seg) INSTA_HIP(hipMemcpyAsync(ts->seg, seg, px, hipMemcpyDeviceToDevice, stream));
    if (depth) INSTA_HIP(hipMemcpyAsync(ts->depth, depth, px * sizeof(float), hipMemcpyDeviceToDevice, stream));
    return INSTA_OK;
}

insta_status train_step(insta_ctx* ctx, const insta_train_config& tc, insta_train_stats* stats, hipStream_t stream) {
    TrainState* ts = ctx->train;
    ts->net_n = 0;  // the step reuses the network plugin's activation buffers
    const uint32_t n_rays = tc.n_rays ? tc.n_rays : ts->rays_per_batch;
    insta_status st = check_train_workspace(ts, n_rays, tc.max_samples, tc.target_samples, "train_step");
    if (st != INSTA_OK) return st;
    // the step evaluates its samples with the training weights (under an Ema
    // optimizer the inference network holds the average between steps)
    const float* inference_src = net_params(ctx);
    if (inference_src != ctx->d_params) {
        ctx->net_src = nullptr;
        st = prepare_network(ctx, stream);
        if (st != INSTA_OK) return st;
    }
    const uint64_t np = ctx->layout.n_params, n_mlp = ctx->layout.n_mlp_params;
    // the marched-sample cap is the step's (a larger reservation changes nothing)
    const uint32_t cap = tc.max_samples;
    // zero gradients with a fill kernel (capturable, unlike memset nodes)
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1024), dim3(256), 0, stream, reinterpret_cast<uint32_t*>(ts->grad), np);
    hipLaunchKernelGGL(train_clear_kernel, dim3(1), dim3(64), 0, stream, ts->counters, cap);
    INSTA_HIP(hipGetLastError());

    Pcg32 rng = Pcg32::seeded(tc.seed);
    rng.advance(ts->rays_total * kRaySeedStride);
    RayArgs ra{};
    ra.box.G = ctx->cfg.grid_size;
    ra.box.dt = ctx->cfg.dt;
    ra.box.max_steps = ctx->cfg.max_steps;
    for (int k = 0; k < 3; ++k) {
        ra.box.lo[k] = ctx->render_lo[k];
        ra.box.hi[k] = ctx->render_hi[k];
    }
    ra.meta = ts->meta;
    ra.grids_lin = ctx->d_grids_lin;
    ra.grid_bytes = (size_t)ctx->cfg.grid_size * ctx->cfg.grid_size * ctx->cfg.grid_size / 8;
    ra.n_images = ts->n_images;
    ra.width = ts->width;
    ra.height = ts->height;
    ra.max_samples = cap;
    ra.fixed_slot = tc.occupancy_slot;
    hipLaunchKernelGGL(train_rays_kernel, dim3((n_rays + 255) / 256), dim3(256), 0, stream, ra, n_rays, ts->rays_total,
                       rng, ts->ray_d, ts->ray_o, ts->ray_frame, ts->ray_img, ts->ray_pix, ts->ray_src, ts->ray_base,
                       ts->ray_cnt, ts->t_buf, ts->ray_T, ts->ray_alive, ts->counters);
    INSTA_HIP(hipGetLastError());
    hipLaunchKernelGGL(train_count_kernel, dim3(1), dim3(64), 0, stream, ts->counters);
    // deform and evaluate the marched samples window by window, only while
    // their rays are still alive
    const dim3 rgrid((n_rays + 255) / 256);
    for (uint32_t w = 0; w < kRounds; ++w) {
        const uint32_t lo = kWin[w], hi = kWin[w + 1];
        hipLaunchKernelGGL(train_gather_kernel, rgrid, dim3(256), 0, stream, lo, hi, ts->ray_base, ts->ray_cnt,
                           ts->ray_alive, ts->ray_gbase, ts->ray_o, ts->ray_d, ts->ray_src, ts->t_buf, n_rays,
                           ts->r_pos, ts->counters, kTRound0 + w);
        INSTA_HIP(hipGetLastError());
        const uint32_t* n_round = ts->counters + kTRound0 + w;
        // the deformation kernels stride over the round's samples: a grid for
        // at most the window's samples and 2^20 threads, not for the sample
        // capacity (empty workgroups of a 2^24-thread grid cost ~15 us a launch)
        const uint32_t round_grid =
            (uint32_t)std::min<uint64_t>(std::min<uint64_t>(cap, (uint64_t)n_rays * (hi - lo)), 1u << 20);
        st = launch_deform_samples(ctx, -1, ts->ray_frame, n_round, round_grid, ts->r_pos, ts->ray_d, ts->s_hit,
                                   ts->r_cpos, ts->r_cdir, ctx->d_misses, stream);
        if (st != INSTA_OK) return st;
        st = launch_network_samples(ctx, n_round, cap, ts->r_cpos, ts->r_cdir, ts->r_out, stream);
        if (st != INSTA_OK) return st;
        hipLaunchKernelGGL(train_round_kernel, rgrid, dim3(256), 0, stream, lo, hi, ctx->cfg.density_logit_bias,
                           ctx->cfg.dt, ts->ray_base, ts->ray_cnt, ts->ray_alive, ts->ray_gbase, ts->ray_T,
                           ts->r_cpos, ts->r_cdir, ts->r_out, ts->s_cpos, ts->s_cdir, ts->s_out, ts->counters);
        INSTA_HIP(hipGetLastError());
    }
    LossArgs la;
    la.images = ts->images;
    la.seg = ts->seg;
    la.depth = ts->depth;
    la.seg_weight = tc.seg_weight;
    la.depth_weight = tc.depth_weight;
    la.width = ts->width;
    la.height = ts->height;
    la.dt = ctx->cfg.dt;
    la.density_bias = ctx->cfg.density_logit_bias;
    la.loss_scale_per_ray = tc.loss_scale / (float)n_rays;
    la.inv_rays = 1.0f / (float)n_rays;
    hipLaunchKernelGGL(train_loss_kernel, dim3((n_rays + 255) / 256), dim3(256), 0, stream, la, rng, ts->ray_src,
                       ts->ray_img, ts->ray_pix, ts->ray_base, ts->ray_cnt, ts->ray_o, ts->ray_d, ts->t_buf,
                       n_rays, ts->s_out,
                       ts->c_idx, ts->c_dout,
                       ts->compact_cap, ts->counters);
    INSTA_HIP(hipGetLastError());
    hipLaunchKernelGGL(train_compact_count_kernel, dim3(1), dim3(64), 0, stream, ts->counters, ts->compact_cap);
    const bool x16 = ctx->cfg.n_extra_dims != 0;
    st = ctx->cfg.n_levels == 8 ? (x16 ? launch_train_forward<8, 16>(ctx, ts, stream) : launch_train_forward<8, 0>(ctx, ts, stream))
                                : (x16 ? launch_train_forward<16, 16>(ctx, ts, stream) : launch_train_forward<16, 0>(ctx, ts, stream));
    if (st != INSTA_OK) return st;
    st = ctx->cfg.n_levels == 8
             ? (x16 ? launch_train_backward<8, 16>(ctx, ts, tc, stream) : launch_train_backward<8, 0>(ctx, ts, tc, stream))
             : (x16 ? launch_train_backward<16, 16>(ctx, ts, tc, stream) : launch_train_backward<16, 0>(ctx, ts, tc, stream));
    if (st != INSTA_OK) return st;
    const uint32_t t = ts->adam_step + (tc.apply_update ? 1u : 0u);
    const float bc1 = tc.apply_update ? 1.0f - (float)std::pow((double)tc.beta1, (double)t) : 1.0f;
    const float bc2 = tc.apply_update ? 1.0f - (float)std::pow((double)tc.beta2, (double)t) : 1.0f;
    hipLaunchKernelGGL(adam_kernel, dim3(2048), dim3(256), 0, stream, ctx->d_params, ts->grad,
                       nullptr, ts->adam_m, ts->adam_v, np, n_mlp,
                       1.0f / tc.loss_scale, tc.learning_rate, tc.beta1, tc.beta2, tc.epsilon, bc1, bc2,
                       tc.l2_reg, tc.apply_update);
    INSTA_HIP(hipGetLastError());
    if (tc.apply_update) {
        if (tc.ema_decay > 0.0f) {
            // the weight average renders use; training keeps its own weights
            ts->ema_step += 1;
            const float debias = (float)(1.0 / (1.0 - std::pow((double)tc.ema_decay, (double)ts->ema_step)));
            hipLaunchKernelGGL(ema_kernel, dim3(2048), dim3(256), 0, stream, ctx->d_params, ts->ema, ts->ema_infer,
                               np, tc.ema_decay, debias, ts->ema_step == 1 ? 1u : 0u);
            INSTA_HIP(hipGetLastError());
            ctx->net_src = ts->ema_infer;
        } else {
            ts->ema_step = 0;
            ctx->net_src = nullptr;
        }
        st = prepare_network(ctx, stream);
        if (st != INSTA_OK) return st;
    } else if (inference_src != net_params(ctx)) {
        ctx->net_src = inference_src == ctx->d_params ? nullptr : inference_src;  // back to what renders used
        st = prepare_network(ctx, stream);
        if (st != INSTA_OK) return st;
    }
    INSTA_HIP(hipMemcpyAsync(ts->h_counters, ts->counters, kTCounters * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             stream));
    INSTA_HIP(hipStreamSynchronize(stream));
    const uint32_t* hc = ts->h_counters;
    if (stats) {
        stats->n_rays = n_rays;
        stats->n_rays_used = hc[kTRays];
        stats->n_samples = hc[kTValid];
        stats->n_samples_used = hc[kTUsed];
        stats->n_samples_evaluated = 0;
        for (uint32_t w = 0; w < kRounds; ++w) stats->n_samples_evaluated += hc[kTRound0 + w];
        float loss;
        std::memcpy(&loss, &hc[kTLoss], 4);
        stats->loss = loss;
    }
    ts->rays_total += n_rays;
    ts->have_grad = true;
    if (tc.apply_update) {
        ts->adam_step = t;
        ts->training_step += 1;
    }
    if (stats) {
        stats->step = ts->training_step;
        stats->adam_step = ts->adam_step;
    }
    // adaptive ray count (testbed_nerf.cu:3402-3410): rays * target /
    // compacted batch, rounded up to a multiple of 128, at most 2^18 rays
    const uint32_t measured = std::max(std::min(hc[kTUsed], ts->compact_cap), 1u);
    double next = (double)n_rays * tc.target_samples / measured;
    next = std::min(std::max(next, 128.0), (double)(1u << 18));
    ts->rays_per_batch = ((uint32_t)next + 127u) & ~127u;
    return INSTA_OK;
}

void reset_inference_average(insta_ctx* ctx) {
    if (ctx->train) ctx->train->ema_step = 0;
    ctx->net_src = nullptr;
}

bool train_has_gradients(const insta_ctx* ctx) { return ctx->train && ctx->train->have_grad; }
bool train_has_set(const insta_ctx* ctx) { return ctx->train && ctx->train->n_images > 0; }
const std::vector<int32_t>& train_frames(const insta_ctx* ctx) { return ctx->train->frames; }
TrainSetView train_set_view(const insta_ctx* ctx) {
    const TrainState* ts = ctx->train;
    if (!ts) return TrainSetView{};
    return TrainSetView{ts->meta, ts->n_images, ts->width, ts->height};
}
uint32_t train_step_count(const insta_ctx* ctx) { return ctx->train ? ctx->train->training_step : 0u; }

insta_status get_gradients(insta_ctx* ctx, float* out, hipStream_t stream) {
    INSTA_HIP(hipMemcpyAsync(out, ctx->train->grad, ctx->layout.n_params * sizeof(float), hipMemcpyDeviceToDevice,
                             stream));
    return INSTA_OK;
}

insta_status train_export(insta_ctx* ctx, bool with_optimizer, TrainSnapshot& out, hipStream_t stream) {
    out = TrainSnapshot{};
    TrainState* ts = ctx->train;
    if (!ts) return INSTA_OK;
    out.present = true;
    out.step = ts->training_step;
    out.adam_step = ts->adam_step;
    out.rays_per_batch = ts->rays_per_batch;
    out.rays_total = ts->rays_total;
    if (ts->n_images && ts->meta) {
        std::vector<TrainImage> meta(ts->n_images);
        INSTA_HIP(hipMemcpyAsync(meta.data(), ts->meta, meta.size() * sizeof(TrainImage), hipMemcpyDeviceToHost,
                                 stream));
        INSTA_HIP(hipStreamSynchronize(stream));
        out.n_images = ts->n_images;
        out.width = ts->width;
        out.height = ts->height;
        out.cams.resize((size_t)ts->n_images * 16);
        for (uint32_t i = 0; i < ts->n_images; ++i) {
            float* c = out.cams.data() + (size_t)i * 16;
            std::memcpy(c, meta[i].M, 12 * sizeof(float));
            c[12] = meta[i].fx;
            c[13] = meta[i].fy;
            c[14] = meta[i].cx;
            c[15] = meta[i].cy;
        }
    }
    if (with_optimizer) {
        const uint64_t np = ctx->layout.n_params;
        out.m.resize(np);
        out.v.resize(np);
        INSTA_HIP(hipMemcpyAsync(out.m.data(), ts->adam_m, np * sizeof(float), hipMemcpyDeviceToHost, stream));
        INSTA_HIP(hipMemcpyAsync(out.v.data(), ts->adam_v, np * sizeof(float), hipMemcpyDeviceToHost, stream));
        INSTA_HIP(hipStreamSynchronize(stream));
    }
    return INSTA_OK;
}

insta_status train_import(insta_ctx* ctx, const TrainSnapshot& in, hipStream_t stream) {
    insta_status st = ensure_train_state(ctx, stream);
    if (st != INSTA_OK) return st;
    TrainState* ts = ctx->train;
    const uint64_t np = ctx->layout.n_params;
    // Testbed::load_snapshot resumes at training_step + 1 (testbed.cu:3011)
    ts->training_step = in.step + 1;
    ts->rays_per_batch = in.rays_per_batch;
    ts->rays_total = in.rays_total;
    if (in.m.size() == np && in.v.size() == np) {
        INSTA_HIP(hipMemcpyAsync(ts->adam_m, in.m.data(), np * sizeof(float), hipMemcpyHostToDevice, stream));
        INSTA_HIP(hipMemcpyAsync(ts->adam_v, in.v.data(), np * sizeof(float), hipMemcpyHostToDevice, stream));
        INSTA_HIP(hipStreamSynchronize(stream));  // pageable sources
        ts->adam_step = in.adam_step;
    } else {
        INSTA_HIP(hipMemsetAsync(ts->adam_m, 0, np * sizeof(float), stream));
        INSTA_HIP(hipMemsetAsync(ts->adam_v, 0, np * sizeof(float), stream));
        ts->adam_step = 0;  // fresh moments restart the bias correction
    }
    return INSTA_OK;
}

insta_status reset_optimizer(insta_ctx* ctx, hipStream_t stream) {
    TrainState* ts = ctx->train;
    const uint64_t np = ctx->layout.n_params;
    // stream-ordered fill kernels: no device-wide stall, capturable
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1024), dim3(256), 0, stream, reinterpret_cast<uint32_t*>(ts->adam_m), np);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1024), dim3(256), 0, stream, reinterpret_cast<uint32_t*>(ts->adam_v), np);
    INSTA_HIP(hipGetLastError());
    ts->adam_step = 0;
    return INSTA_OK;
}

// ---- NerfNetwork plugin entry points --------------------------------------
namespace {
insta_status ensure_net_workspace(insta_ctx* ctx, uint32_t n, hipStream_t) {
    if (!ctx->train) {
        set_error("network plugin: no workspace: reserve one with insta_reserve (train_samples, train_batch)");
        return INSTA_ERR_STATE;
    }
    TrainState* ts = ctx->train;
    insta_status st = check_train_workspace(ts, 0, n, n, "network plugin");
    if (st != INSTA_OK) return st;
    if (ctx->cfg.n_extra_dims && ts->net_extra_cap < n) {
        set_error("network plugin: extra-dims workspace smaller than the batch: reserve with insta_reserve");
        return INSTA_ERR_STATE;
    }
    return INSTA_OK;
}
}  // namespace

insta_status reserve_training(insta_ctx* ctx, uint32_t rays, uint32_t samples, uint32_t batch, hipStream_t stream) {
    insta_status st = ensure_train_state(ctx, stream);
    if (st != INSTA_OK) return st;
    TrainState* ts = ctx->train;
    st = grow_train_workspace(ctx, ts, rays, samples, batch);
    if (st != INSTA_OK) return st;
    const uint32_t X = ctx->cfg.n_extra_dims;
    if (X && ts->net_extra_cap < ts->compact_cap) {
        dev_free(ts->net_extra);
        ts->net_extra_cap = 0;
        if (hipMalloc(reinterpret_cast<void**>(&ts->net_extra), (size_t)ts->compact_cap * X * sizeof(float)) !=
            hipSuccess) {
            set_error("network plugin workspace allocation failed");
            return INSTA_ERR_OOM;
        }
        ts->net_extra_cap = ts->compact_cap;
    }
    return INSTA_OK;
}

insta_status network_forward(insta_ctx* ctx, const float* coords, uint32_t n, uint32_t stride, float* out,
                             hipStream_t stream) {
    insta_status st = ensure_net_workspace(ctx, n, stream);
    if (st != INSTA_OK) return st;
    TrainState* ts = ctx->train;
    const uint32_t X = ctx->cfg.n_extra_dims;
    hipLaunchKernelGGL(net_prepare_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, coords, n, stride, X, ts->s_cpos,
                       ts->s_cdir, ts->c_idx, ts->net_extra, ts->counters);
    INSTA_HIP(hipGetLastError());
    float4* o = reinterpret_cast<float4*>(out);
    const bool x16 = X != 0;
    st = ctx->cfg.n_levels == 8 ? (x16 ? launch_net_forward<8, 16>(ctx, ts, o, stream) : launch_net_forward<8, 0>(ctx, ts, o, stream))
                                : (x16 ? launch_net_forward<16, 16>(ctx, ts, o, stream) : launch_net_forward<16, 0>(ctx, ts, o, stream));
    if (st != INSTA_OK) return st;
    ts->net_n = n;
    return INSTA_OK;
}

uint32_t network_forward_batch(const insta_ctx* ctx) { return ctx->train ? ctx->train->net_n : 0u; }

insta_status network_backward(insta_ctx* ctx, uint32_t n, uint32_t stride, const float* dL_dout, float* grad_out,
                              bool accumulate, float* dL_dinput, hipStream_t stream) {
    TrainState* ts = ctx->train;
    const uint64_t np = ctx->layout.n_params;
    INSTA_HIP(hipMemcpyAsync(ts->c_dout, dL_dout, (size_t)n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1024), dim3(256), 0, stream, reinterpret_cast<uint32_t*>(ts->grad), np);
    INSTA_HIP(hipGetLastError());
    insta_train_config tc{};
    tc.grid_grad_fp16 = 0;
    const bool x16 = ctx->cfg.n_extra_dims != 0;
    insta_status st =
        ctx->cfg.n_levels == 8
            ? (x16 ? launch_train_backward<8, 16>(ctx, ts, tc, stream) : launch_train_backward<8, 0>(ctx, ts, tc, stream))
            : (x16 ? launch_train_backward<16, 16>(ctx, ts, tc, stream) : launch_train_backward<16, 0>(ctx, ts, tc, stream));
    if (st != INSTA_OK) return st;
    hipLaunchKernelGGL(net_grad_out_kernel, dim3(1024), dim3(256), 0, stream, ts->grad, grad_out, np,
                       accumulate ? 1u : 0u);
    INSTA_HIP(hipGetLastError());
    if (dL_dinput) {
        st = ctx->cfg.n_levels == 8
                 ? (x16 ? launch_net_input_grad<8, 16>(ctx, ts, n, dL_dinput, stride, stream)
                        : launch_net_input_grad<8, 0>(ctx, ts, n, dL_dinput, stride, stream))
                 : (x16 ? launch_net_input_grad<16, 16>(ctx, ts, n, dL_dinput, stride, stream)
                        : launch_net_input_grad<16, 0>(ctx, ts, n, dL_dinput, stride, stream));
        if (st != INSTA_OK) return st;
    }
    return INSTA_OK;
}

}  // namespace insta
