// network_mfma.hip -- fused hash encoding + density/colour MLPs on MFMA.
//
// One workgroup (4 waves) loops over tiles of 256 samples:
//   phase A  one thread per sample: hash-grid encoding (same arithmetic as
//            the scalar path, so FP32 features are bit-identical) written to
//            LDS (the FP16/BF16 kernel also stores the 16 SH values there)
//   phase B  each wave runs its 64 samples (4 column blocks of 16) through
//            the MLPs with 16x16 MFMA tiles: weights are the A operand
//            (rows = output neurons), samples the B operand (columns).
// A layer's 16x16 accumulator holds, in lane l, rows 4(l/16)..4(l/16)+3 of
// column l%16 -- exactly the B-operand rows the next layer needs when its
// K dimension is visited in the order {16t + 4g + i}; the weight fragments
// are pre-arranged in that order (prep kernels below), so activations never
// leave registers between layers.
//   FP32  v_mfma_f32_16x16x4f32  (K blocks of 4; weight fragments read per
//         use through the vector cache, SH evaluated per lane in phase B)
//   FP16  v_mfma_f32_16x16x32_f16, BF16 v_mfma_f32_16x16x32_bf16 (K blocks
//         of 32 = two 16-row accumulators; weights held in VGPRs), hash
//         tables stored FP16, accumulation FP32
// Shapes: width 64, density 1 hidden layer -> 16 outputs, colour input 16 +
// 16 SH, 2 hidden layers -> 3 outputs (the workload MLPs).
// Extra dims (16 expression values appended to the density input): a
// sample's 16 values are either all 1.0 or its frame's row, so their share of
// the first layer is one of n_frames + 1 vectors W0[:, enc:] x, computed in
// FP32 when the weights or a frame's values change (extra_bias_kernel) and
// loaded as the first layer's starting accumulator; the per-sample selector
// (0: ones, 1 + f: frame f) rides in the canonical direction's w.
#include "context.hpp"

namespace insta {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int kTile = 256;
constexpr int kW = 64;
constexpr int kDOut = 16;

// ---------------------------------------------------------------------------
// weight fragment layout
//   FP32: fragment = one K block of 4 for one 16-row output tile; lane l
//         holds W[row0 + l%16][k(l/16)].  Order: L0 (4 tiles x DIN/4),
//         density out (16 blocks), rgb L0 (4 tiles x 8), rgb L1 (4 x 16),
//         rgb out (16).
//   FP16/BF16: fragment = one K block of 32 for one output tile; lane l holds
//         8 values W[row0 + l%16][k(l/16, j)], j = 0..7.  Order: L0 (4 tiles
//         x ceil(DIN/32)), density out (2), rgb L0 (4), rgb L1 (4 x 2),
//         rgb out (2).
// k for chained layers, K block kb of 32 (two source tiles 2kb, 2kb+1):
//   j < 4: 32 kb + 4g + j,  j >= 4: 32 kb + 16 + 4g + (j - 4)
// k for the colour input (density outputs 0..15, SH 16..31):
//   j < 4: 4g + j (density),  j >= 4: 16 + 4g + (j - 4) (SH)
// ---------------------------------------------------------------------------
struct MlpOffsets {
    uint32_t w0, w1, r0, r1, r2;  // float offsets of each matrix in params
    uint32_t din;                 // W0 row length: encoding + extra dims
    uint32_t enc;                 // encoding columns (the MFMA part of W0)
};

MlpOffsets mlp_offsets(const insta_ctx* ctx) {
    MlpOffsets o;
    const uint32_t din = ctx->layout.density_in;
    o.din = din;
    o.enc = ctx->cfg.n_levels * ctx->cfg.n_features_per_level;
    o.w0 = 0;
    o.w1 = o.w0 + kW * din;
    o.r0 = o.w1 + kDOut * kW;
    o.r1 = o.r0 + kW * 32;
    o.r2 = o.r1 + kW * kW;
    return o;
}

__device__ inline float wget(const float* p, uint32_t off, uint32_t rows, uint32_t cols, uint32_t r, uint32_t c) {
    return (r < rows && c < cols) ? p[off + r * cols + c] : 0.0f;
}

// FP32 fragments: index f, lane l -> value
__global__ void prep_f32_kernel(const float* __restrict__ p, MlpOffsets o, float* __restrict__ frag, uint32_t n_frag) {
    const uint32_t f = blockIdx.x, l = threadIdx.x;
    if (f >= n_frag) return;
    const uint32_t m = l % 16, g = l / 16;
    const uint32_t kb0 = (o.enc + 3) / 4;
    uint32_t idx = f;
    float v = 0.0f;
    if (idx < 4 * kb0) {  // density L0: tile t, block kb (natural k order)
        const uint32_t t = idx / kb0, kb = idx % kb0;
        v = wget(p, o.w0, kW, o.din, 16 * t + m, 4 * kb + g < o.enc ? 4 * kb + g : o.din);
    } else if ((idx -= 4 * kb0) < 16) {  // density out: block (t = idx/4, i = idx%4)
        const uint32_t t = idx / 4, i = idx % 4;
        v = wget(p, o.w1, kDOut, kW, m, 16 * t + 4 * g + i);
    } else if ((idx -= 16) < 32) {  // rgb L0: tile t, block b: b<4 density part, b>=4 SH part
        const uint32_t t = idx / 8, b = idx % 8;
        const uint32_t k = b < 4 ? 4 * g + b : 16 + 4 * g + (b - 4);
        v = wget(p, o.r0, kW, 32, 16 * t + m, k);
    } else if ((idx -= 32) < 64) {  // rgb L1: tile t, block (s = b/4 source tile, i = b%4)
        const uint32_t t = idx / 16, b = idx % 16;
        v = wget(p, o.r1, kW, kW, 16 * t + m, 16 * (b / 4) + 4 * g + (b % 4));
    } else {  // rgb out
        idx -= 64;
        v = wget(p, o.r2, 3, kW, m, 16 * (idx / 4) + 4 * g + (idx % 4));
    }
    frag[(size_t)f * 64 + l] = v;
}

template <typename T>
__global__ void prep_x32_kernel(const float* __restrict__ p, MlpOffsets o, T* __restrict__ frag, uint32_t n_frag) {
    const uint32_t f = blockIdx.x, l = threadIdx.x;
    if (f >= n_frag) return;
    const uint32_t m = l % 16, g = l / 16;
    const uint32_t kb0 = (o.enc + 31) / 32;
    for (uint32_t j = 0; j < 8; ++j) {
        uint32_t idx = f;
        float v;
        const uint32_t kc = j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4);  // chained k within a 32 block
        if (idx < 4 * kb0) {
            const uint32_t t = idx / kb0, kb = idx % kb0, k = 32 * kb + 8 * g + j;
            v = wget(p, o.w0, kW, o.din, 16 * t + m, k < o.enc ? k : o.din);
        } else if ((idx -= 4 * kb0) < 2) {
            v = wget(p, o.w1, kDOut, kW, m, 32 * idx + kc);
        } else if ((idx -= 2) < 4) {
            v = wget(p, o.r0, kW, 32, 16 * idx + m, kc);
        } else if ((idx -= 4) < 8) {
            v = wget(p, o.r1, kW, kW, 16 * (idx / 2) + m, 32 * (idx % 2) + kc);
        } else {
            idx -= 8;
            v = wget(p, o.r2, 3, kW, m, 32 * idx + kc);
        }
        frag[((size_t)f * 64 + l) * 8 + j] = (T)v;
    }
}

constexpr uint32_t kFragX32 = 4 + 2 + 4 + 8 + 2;  // with DIN <= 32

// rows [first, first + n) of the extra-dims table: row 0 = W0[:, enc:] . 1,
// row 1 + f = W0[:, enc:] . (frame f's values); one thread per output neuron
__global__ void extra_bias_kernel(const float* __restrict__ p, MlpOffsets o, uint32_t n_extra,
                                  const float* __restrict__ frame_extra, uint32_t first, float* __restrict__ bias) {
    const uint32_t r = first + blockIdx.x, j = threadIdx.x;
    const float* w = p + o.w0 + (size_t)j * o.din + o.enc;
    const float* x = r ? frame_extra + (size_t)(r - 1) * n_extra : nullptr;
    float acc = 0.0f;
    for (uint32_t e = 0; e < n_extra; ++e) acc = __builtin_fmaf(w[e], x ? x[e] : 1.0f, acc);
    bias[(size_t)r * kW + j] = acc;
}

// frame values from the first n_extra expression coefficients (zero-padded)
__global__ void extra_from_coeffs_kernel(const float* __restrict__ coeffs, uint32_t n_coeffs, uint32_t n_extra,
                                         float* __restrict__ dst) {
    const uint32_t e = threadIdx.x;
    if (e < n_extra) dst[e] = e < n_coeffs ? coeffs[e] : 0.0f;
}

__global__ void table_to_half_kernel(const float* __restrict__ src, _Float16* __restrict__ dst, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (_Float16)src[i];
}

// ---------------------------------------------------------------------------
// sample sources
// ---------------------------------------------------------------------------
struct SoaSrc {
    const float4* cpos;
    const float4* cdir;
    float4* out;
    const uint32_t* n_dev;
    INSTA_DEV uint32_t count() const { return *n_dev; }
    INSTA_DEV void load(uint32_t i, V3& p, V3& d) const {
        const float4 a = cpos[i], b = cdir[i];
        p = v3(a.x, a.y, a.z);
        d = v3(b.x, b.y, b.z);
    }
    INSTA_DEV uint32_t sel(uint32_t i) const { return __float_as_uint(cdir[i].w); }
};
struct AosSrc {
    const float* coords;
    uint32_t stride, n;
    float4* out;
    INSTA_DEV uint32_t count() const { return n; }
    INSTA_DEV void load(uint32_t i, V3& p, V3& d) const {
        const float* c = coords + (size_t)i * stride;
        p = v3(c[0], c[1], c[2]);
        d = v3(c[8], c[9], c[10]);
    }
    INSTA_DEV uint32_t sel(uint32_t) const { return 0u; }  // extra dims take the scalar kernel
};

INSTA_DEV void sh16_dev(float x, float y, float z, float* o) {
    const float xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
    o[0] = 0.28209479177387814f;
    o[1] = -0.48860251190291987f * y;
    o[2] = 0.48860251190291987f * z;
    o[3] = -0.48860251190291987f * x;
    o[4] = 1.0925484305920792f * xy;
    o[5] = -1.0925484305920792f * yz;
    o[6] = 0.94617469575755997f * z2 - 0.31539156525251999f;
    o[7] = -1.0925484305920792f * xz;
    o[8] = 0.54627421529603959f * x2 - 0.54627421529603959f * y2;
    o[9] = 0.59004358992664352f * y * (-3.0f * x2 + y2);
    o[10] = 2.8906114426405538f * xy * z;
    o[11] = 0.45704579946446572f * y * (1.0f - 5.0f * z2);
    o[12] = 0.3731763325901154f * z * (5.0f * z2 - 3.0f);
    o[13] = 0.45704579946446572f * x * (1.0f - 5.0f * z2);
    o[14] = 1.4453057213202769f * z * (x2 - y2);
    o[15] = 0.59004358992664352f * x * (-x2 + 3.0f * y2);
}

// one level of the hash grid, FP32 or FP16 table, FP32 interpolation (the
// same operation order as the scalar encode)
template <typename TT>
INSTA_DEV float2 encode_level2(const TT* __restrict__ table, const GridParams& gp, uint32_t l, V3 x) {
    const float sc = gp.scale[l];
    const float px = x.x * sc + 0.5f, py = x.y * sc + 0.5f, pz = x.z * sc + 0.5f;
    const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    const float frac[3] = {px - fx, py - fy, pz - fz};
    const uint32_t g[3] = {(uint32_t)(int32_t)fx, (uint32_t)(int32_t)fy, (uint32_t)(int32_t)fz};
    const bool dense = (gp.dense_mask >> l) & 1u;
    const uint32_t res = gp.res[l], size = gp.size[l];
    const TT* base = table + (size_t)gp.offset[l] * 2;
    float a0 = 0.0f, a1 = 0.0f;
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
        uint32_t idx;
        if (dense) {
            idx = cg[0] + cg[1] * res + cg[2] * (res * res);
            if (idx >= size) idx %= size;
        } else {
            idx = (cg[0] ^ (cg[1] * 2654435761u) ^ (cg[2] * 805459861u)) & (size - 1u);
        }
        float v0, v1;
        if constexpr (sizeof(TT) == 4) {
            const float2 v = *reinterpret_cast<const float2*>(base + (size_t)idx * 2);
            v0 = v.x;
            v1 = v.y;
        } else {
            typedef _Float16 h2 __attribute__((ext_vector_type(2)));
            const h2 v = *reinterpret_cast<const h2*>(base + (size_t)idx * 2);
            v0 = (float)v.x;
            v1 = (float)v.y;
        }
        a0 = a0 + w * v0;
        a1 = a1 + w * v1;
    }
    return make_float2(a0, a1);
}

// ---------------------------------------------------------------------------
// FP32 MFMA kernel
// ---------------------------------------------------------------------------
constexpr int kF32Stride = 33;  // floats per sample in LDS: 32 features + 1 pad

// XCD-aware tile order for the persistent grid.  Workgroups are dispatched
// round-robin over the 8 XCDs (block b on XCD b % 8), each with its own L2:
// the blocks of one XCD walk one contiguous eighth of the samples, so an
// XCD's L2 holds the table entries of one region of the image rather than
// of all of it.  Grids that are not a multiple of 8 blocks stride plainly.
struct XcdTiles {
    uint32_t tile, end, step;
    INSTA_DEV explicit XcdTiles(uint32_t n) {
        const uint32_t tiles = (n + kTile - 1) / kTile;
        if (gridDim.x % 8u == 0) {
            const uint32_t xcd = blockIdx.x % 8u, chunk = (tiles + 7u) / 8u;
            tile = xcd * chunk + blockIdx.x / 8u;
            end = min(tiles, (xcd + 1u) * chunk);
            step = gridDim.x / 8u;
        } else {
            tile = blockIdx.x;
            end = tiles;
            step = gridDim.x;
        }
    }
};

template <int L, int X, class Src>
__global__ __launch_bounds__(256) void network_f32_kernel(const float* __restrict__ table, GridParams gp,
                                                          const float* __restrict__ frag,
                                                          const float* __restrict__ xbias, uint32_t xrows,
                                                          Src src) {
    constexpr int DIN = L * 2;
    static_assert(DIN < kF32Stride, "LDS row holds the features and the extra-dims selector");
    constexpr int KB0 = DIN / 4;
    __shared__ float s_x[kTile * kF32Stride];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t n = src.count();
    const uint32_t g = lane >> 4, col = lane & 15u;
    for (XcdTiles it(n); it.tile < it.end; it.tile += it.step) {
        const uint32_t base = it.tile * kTile;
        __syncthreads();
        {
            const uint32_t i = base + tid;
            float* row = s_x + tid * kF32Stride;
            if (i < n) {
                V3 p, d;
                src.load(i, p, d);
#pragma unroll 2
                for (int l = 0; l < L; ++l) {
                    const float2 e = encode_level2<float>(table, gp, (uint32_t)l, p);
                    row[2 * l] = e.x;
                    row[2 * l + 1] = e.y;
                }
            } else {
                for (int k = 0; k < DIN; ++k) row[k] = 0.0f;
            }
            if constexpr (X > 0) {
                const uint32_t sel = i < n ? src.sel(i) : 0u;
                row[kF32Stride - 1] = __uint_as_float(sel < xrows ? sel : 0u);  // rows of xbias only
            }
        }
        __syncthreads();
        for (uint32_t nb = 0; nb < 4; ++nb) {
            const uint32_t s = wave * 64 + nb * 16 + col;  // sample (column) of this lane
            const uint32_t i = base + s;
            const float* xr = s_x + s * kF32Stride;
            // SH of this column's direction; lane group g feeds rows 4g..4g+3
            float sh[16];
            {
                V3 p, d = v3(0.0f, 0.0f, 1.0f);
                if (i < n) src.load(i, p, d);
                sh16_dev(d.x * 2.0f - 1.0f, d.y * 2.0f - 1.0f, d.z * 2.0f - 1.0f, sh);
            }
            float shg[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) shg[b] = g == 0 ? sh[b] : (g == 1 ? sh[4 + b] : (g == 2 ? sh[8 + b] : sh[12 + b]));
            const float* W = frag + lane;
            const float* xb = xbias;
            if constexpr (X > 0) xb += (size_t)__float_as_uint(xr[kF32Stride - 1]) * kW + 4 * g;
            f4 h[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                if constexpr (X > 0) acc = *reinterpret_cast<const f4*>(xb + 16 * t);
#pragma unroll
                for (int kb = 0; kb < KB0; ++kb)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W[(t * KB0 + kb) * 64], xr[4 * kb + g], acc, 0, 0, 0);
                h[t] = acc;
            }
            W += 4 * KB0 * 64;
            f4 dn = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    dn = __builtin_amdgcn_mfma_f32_16x16x4f32(W[(t * 4 + b) * 64], fmaxf(h[t][b], 0.0f), dn, 0, 0, 0);
            W += 16 * 64;
            f4 r[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W[(t * 8 + b) * 64], dn[b], acc, 0, 0, 0);
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W[(t * 8 + 4 + b) * 64], shg[b], acc, 0, 0, 0);
                r[t] = acc;
            }
            W += 32 * 64;
            f4 q[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(W[(t * 16 + s2 * 4 + b) * 64], fmaxf(r[s2][b], 0.0f),
                                                                   acc, 0, 0, 0);
                q[t] = acc;
            }
            W += 64 * 64;
            f4 o = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    o = __builtin_amdgcn_mfma_f32_16x16x4f32(W[(s2 * 4 + b) * 64], fmaxf(q[s2][b], 0.0f), o, 0, 0, 0);
            if (g == 0 && i < n) src.out[i] = make_float4(o[0], o[1], o[2], dn[0]);
        }
    }
}

// ---------------------------------------------------------------------------
// FP16 / BF16 MFMA kernel (16x16x32), FP16 hash tables
// ---------------------------------------------------------------------------
constexpr int kX32Stride = 56;  // halfs per sample in LDS: 32 features + 16 SH + 8 pad

template <typename T>
struct Vec8;
template <>
struct Vec8<_Float16> {
    typedef h8 type;
    static INSTA_DEV f4 mfma(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <>
struct Vec8<__bf16> {
    typedef b8 type;
    static INSTA_DEV f4 mfma(b8 a, b8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

template <typename T>
INSTA_DEV typename Vec8<T>::type pack_relu(f4 a, f4 b) {
    typename Vec8<T>::type v;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i] = (T)fmaxf(a[i], 0.0f);
        v[4 + i] = (T)fmaxf(b[i], 0.0f);
    }
    return v;
}

template <int L, int X, typename T, class Src>
__global__ __launch_bounds__(256, 4) void network_x32_kernel(const _Float16* __restrict__ table, GridParams gp,
                                                          const T* __restrict__ frag, const float* __restrict__ xbias,
                                                          uint32_t xrows, Src src) {
    typedef typename Vec8<T>::type V8;
    constexpr int DIN = L * 2;
    static_assert(DIN <= 32, "one K block for the first layer");
    __shared__ __attribute__((aligned(16))) T s_x[kTile * kX32Stride];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t g = lane >> 4, col = lane & 15u;
    // weight fragments into registers (kFragX32 x 8 values per lane)
    V8 A[kFragX32];
#pragma unroll
    for (uint32_t f = 0; f < kFragX32; ++f) A[f] = *reinterpret_cast<const V8*>(frag + ((size_t)f * 64 + lane) * 8);
    const uint32_t n = src.count();
    for (XcdTiles it(n); it.tile < it.end; it.tile += it.step) {
        const uint32_t base = it.tile * kTile;
        __syncthreads();
        {
            const uint32_t i = base + tid;
            T* row = s_x + tid * kX32Stride;
            if (i < n) {
                V3 p, d;
                src.load(i, p, d);
#pragma unroll 2
                for (int l = 0; l < L; ++l) {
                    const float2 e = encode_level2<_Float16>(table, gp, (uint32_t)l, p);
                    row[2 * l] = (T)e.x;
                    row[2 * l + 1] = (T)e.y;
                }
#pragma unroll
                for (int k = DIN; k < 32; ++k) row[k] = (T)0.0f;
                float sh[16];
                sh16_dev(d.x * 2.0f - 1.0f, d.y * 2.0f - 1.0f, d.z * 2.0f - 1.0f, sh);
#pragma unroll
                for (int k = 0; k < 16; ++k) row[32 + k] = (T)sh[k];
            } else {
                for (int k = 0; k < 48; ++k) row[k] = (T)0.0f;
            }
            // extra-dims selector in the row's padding (halfs 48-49, 4-byte aligned)
            if constexpr (X > 0) {
                const uint32_t sel = i < n ? src.sel(i) : 0u;
                *reinterpret_cast<uint32_t*>(row + 48) = sel < xrows ? sel : 0u;  // rows of xbias only
            }
        }
        __syncthreads();
        for (uint32_t nb = 0; nb < 4; ++nb) {
            const uint32_t s = wave * 64 + nb * 16 + col;
            const T* xr = s_x + s * kX32Stride;
            V8 bx;
#pragma unroll
            for (int j = 0; j < 8; ++j) bx[j] = xr[8 * g + j];
            f4 h[4];
            if constexpr (X > 0) {
                const float* xb = xbias + (size_t)*reinterpret_cast<const uint32_t*>(xr + 48) * kW + 4 * g;
#pragma unroll
                for (int t = 0; t < 4; ++t) h[t] = Vec8<T>::mfma(A[t], bx, *reinterpret_cast<const f4*>(xb + 16 * t));
            } else {
#pragma unroll
                for (int t = 0; t < 4; ++t) h[t] = Vec8<T>::mfma(A[t], bx, f4{0.0f, 0.0f, 0.0f, 0.0f});
            }
            f4 dn = Vec8<T>::mfma(A[4], pack_relu<T>(h[0], h[1]), f4{0.0f, 0.0f, 0.0f, 0.0f});
            dn = Vec8<T>::mfma(A[5], pack_relu<T>(h[2], h[3]), dn);
            V8 rin;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                rin[i] = (T)dn[i];
                rin[4 + i] = xr[32 + 4 * g + i];
            }
            f4 r[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) r[t] = Vec8<T>::mfma(A[6 + t], rin, f4{0.0f, 0.0f, 0.0f, 0.0f});
            const V8 r01 = pack_relu<T>(r[0], r[1]), r23 = pack_relu<T>(r[2], r[3]);
            f4 q[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                q[t] = Vec8<T>::mfma(A[10 + 2 * t], r01, f4{0.0f, 0.0f, 0.0f, 0.0f});
                q[t] = Vec8<T>::mfma(A[11 + 2 * t], r23, q[t]);
            }
            f4 o = Vec8<T>::mfma(A[18], pack_relu<T>(q[0], q[1]), f4{0.0f, 0.0f, 0.0f, 0.0f});
            o = Vec8<T>::mfma(A[19], pack_relu<T>(q[2], q[3]), o);
            const uint32_t i = base + s;
            if (g == 0 && i < n) src.out[i] = make_float4(o[0], o[1], o[2], dn[0]);
        }
    }
}

uint32_t grid_for(uint32_t n_max) {
    const uint32_t tiles = (n_max + kTile - 1) / kTile;
    const uint32_t cap = 256 * 8;
    if (tiles >= cap) return cap;
    return tiles >= 8 ? (tiles + 7u) / 8u * 8u : (tiles ? tiles : 1);  // whole XCD rounds when possible
}

}  // namespace

bool mfma_network_supported(const insta_config& c) {
    return c.precision != INSTA_PRECISION_FP32_SCALAR && c.n_neurons == 64 && c.density_out == 16 && c.sh_degree == 4 &&
           (c.n_extra_dims == 0 || c.n_extra_dims == 16) &&
           c.density_hidden_layers == 1 && c.rgb_hidden_layers == 2 && c.n_features_per_level == 2 &&
           (c.n_levels == 8 || c.n_levels == 16);
}

insta_status update_extra_rows(insta_ctx* ctx, uint32_t first, uint32_t n, hipStream_t stream) {
    if (!ctx->d_extra_bias || n == 0) return INSTA_OK;
    hipLaunchKernelGGL(extra_bias_kernel, dim3(n), dim3(kW), 0, stream, net_params(ctx), mlp_offsets(ctx),
                       ctx->cfg.n_extra_dims, ctx->d_frame_extra, first, ctx->d_extra_bias);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_extra_from_coeffs(insta_ctx* ctx, uint32_t frame, const float* coeffs, uint32_t n_coeffs,
                                      hipStream_t stream) {
    if (!ctx->d_frame_extra) return INSTA_OK;
    const uint32_t X = ctx->cfg.n_extra_dims;
    hipLaunchKernelGGL(extra_from_coeffs_kernel, dim3(1), dim3(64), 0, stream, coeffs, n_coeffs, X,
                       ctx->d_frame_extra + (size_t)frame * X);
    INSTA_HIP(hipGetLastError());
    return update_extra_rows(ctx, frame + 1, 1, stream);
}

// Called after the FP32 parameters change: arrange weight fragments,
// (reduced precision) convert the hash tables, refresh the extra-dims table.
insta_status prepare_network(insta_ctx* ctx, hipStream_t stream) {
    insta_status st = update_extra_rows(ctx, 0, ctx->n_slots, stream);
    if (st != INSTA_OK) return st;
    if (!mfma_network_supported(ctx->cfg)) return INSTA_OK;
    const MlpOffsets o = mlp_offsets(ctx);
    if (ctx->cfg.precision == INSTA_PRECISION_FP32) {
        const uint32_t nfrag = 4 * ((o.enc + 3) / 4) + 16 + 32 + 64 + 16;
        hipLaunchKernelGGL(prep_f32_kernel, dim3(nfrag), dim3(64), 0, stream, net_params(ctx), o,
                           static_cast<float*>(ctx->d_wfrag), nfrag);
    } else {
        if (ctx->cfg.precision == INSTA_PRECISION_FP16)
            hipLaunchKernelGGL(prep_x32_kernel<_Float16>, dim3(kFragX32), dim3(64), 0, stream, net_params(ctx), o,
                               static_cast<_Float16*>(ctx->d_wfrag), kFragX32);
        else
            hipLaunchKernelGGL(prep_x32_kernel<__bf16>, dim3(kFragX32), dim3(64), 0, stream, net_params(ctx), o,
                               static_cast<__bf16*>(ctx->d_wfrag), kFragX32);
        const uint64_t nt = ctx->layout.n_grid_params;
        hipLaunchKernelGGL(table_to_half_kernel, dim3((uint32_t)((nt + 255) / 256)), dim3(256), 0, stream,
                           net_params(ctx) + ctx->layout.n_mlp_params, static_cast<_Float16*>(ctx->d_table_h), nt);
    }
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

template <int X, class Src>
static insta_status launch_mfma_x(insta_ctx* ctx, Src src, uint32_t n_max, hipStream_t stream) {
    const GridParams gp = grid_params(ctx);
    const dim3 grid(grid_for(n_max)), block(256);
    const float* table = net_params(ctx) + ctx->layout.n_mlp_params;
    const _Float16* table_h = static_cast<const _Float16*>(ctx->d_table_h);
    const float* xb = ctx->d_extra_bias;
    const uint32_t xr = ctx->n_slots;  // rows of the extra-dims table
    const uint32_t L = ctx->cfg.n_levels;
    switch (ctx->cfg.precision) {
        case INSTA_PRECISION_FP32: {
            const float* fr = static_cast<const float*>(ctx->d_wfrag);
            if (L == 8) hipLaunchKernelGGL((network_f32_kernel<8, X, Src>), grid, block, 0, stream, table, gp, fr, xb, xr, src);
            else hipLaunchKernelGGL((network_f32_kernel<16, X, Src>), grid, block, 0, stream, table, gp, fr, xb, xr, src);
            break;
        }
        case INSTA_PRECISION_FP16: {
            const _Float16* fr = static_cast<const _Float16*>(ctx->d_wfrag);
            if (L == 8) hipLaunchKernelGGL((network_x32_kernel<8, X, _Float16, Src>), grid, block, 0, stream, table_h, gp, fr, xb, xr, src);
            else hipLaunchKernelGGL((network_x32_kernel<16, X, _Float16, Src>), grid, block, 0, stream, table_h, gp, fr, xb, xr, src);
            break;
        }
        default: {
            const __bf16* fr = static_cast<const __bf16*>(ctx->d_wfrag);
            if (L == 8) hipLaunchKernelGGL((network_x32_kernel<8, X, __bf16, Src>), grid, block, 0, stream, table_h, gp, fr, xb, xr, src);
            else hipLaunchKernelGGL((network_x32_kernel<16, X, __bf16, Src>), grid, block, 0, stream, table_h, gp, fr, xb, xr, src);
            break;
        }
    }
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_network_mfma_soa(insta_ctx* ctx, const uint32_t* n_dev, uint32_t n_max, const float4* cpos,
                                     const float4* cdir, float4* out, hipStream_t stream) {
    if (n_max == 0) return INSTA_OK;
    SoaSrc s{cpos, cdir, out, n_dev};
    return ctx->cfg.n_extra_dims ? launch_mfma_x<16>(ctx, s, n_max, stream) : launch_mfma_x<0>(ctx, s, n_max, stream);
}

insta_status launch_network_mfma_aos(insta_ctx* ctx, const float* coords, uint32_t n, uint32_t stride, float* out,
                                     hipStream_t stream) {
    if (n == 0) return INSTA_OK;
    if (ctx->cfg.n_extra_dims) {
        set_error("network_inference: extra dims take the per-sample FP32 kernel");
        return INSTA_ERR_UNSUPPORTED;
    }
    AosSrc s{coords, stride, n, reinterpret_cast<float4*>(out)};
    return launch_mfma_x<0>(ctx, s, n, stream);
}

}  // namespace insta
