// bvh.hip -- per-frame mesh preparation on the device: triangles, BVH, frames.
//
// Replaces the reference's per-frame host work (src/tiny_mesh.cu:134-178:
// build_triangles + TriangleBvh::build on the CPU, then an H2D copy) with one
// workgroup per frame:
//   1. centroids (float, (a+b+c)/3) and their AABB (workgroup reduction);
//   2. per axis, keys (18-bit quantised centroid << 14 | face id) sorted by an
//      LDS bitonic sort (<= 16,384 keys), giving three presorted id lists;
//   3. for every binary depth of the implicit 4-ary heap, each segment picks
//      its longest-extent axis and all three lists are stable-partitioned at
//      the segment's count median (block-wide scans) -- a median-split tree
//      like triangle_bvh.cu:548-631, but exactly reproducible;
//   4. BVH-order triangle records, leaf boxes over the 4^k count-balanced
//      leaves, inner boxes bottom-up;
//   5. per-triangle frames tri2projection (include/rta/affine.cuh:26-39) and
//      their inverses, built and inverted in FP64, stored FP32.
// Several frames build concurrently, one per CU.
#include "context.hpp"

namespace insta {
namespace {

constexpr uint32_t kBuildThreads = 1024;

struct BuildArgs {
    const float* verts;        // [n][nv*3] input vertices for the slots being built
    const float* canon_verts;  // canonical vertices (slot 0 store) for area ratios
    const int32_t* faces;
    const uint32_t* masks;
    float* verts_store;
    float4* nodes;
    float4* btris;
    float4* proj;
    float4* proj_inv;
    float4* cen;
    uint16_t* lists;           // [slot][2 buffers][3 axes][tri_cap] build scratch
    uint32_t* qcoord;          // [slot][3][tri_cap] quantised centroids
    uint32_t first_slot, nv, nf, k, leaf_base, node_cap, tri_cap, vert_cap;
};

struct D3 {
    double x, y, z;
};
__device__ inline D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline D3 dnormalized(D3 a) {
    double s = ddot(a, a);
    if (s > 0.0) {
        double n = sqrt(s);
        return {a.x / n, a.y / n, a.z / n};
    }
    return a;
}

__device__ inline V3 load_v(const float* V, int32_t i) { return v3(V[3 * i], V[3 * i + 1], V[3 * i + 2]); }

__device__ inline uint32_t quantise18(float c, float lo, float e) {
    if (!(e > 0.0f)) return 0u;
    float q = (c - lo) / e * 262144.0f;
    if (!(q > 0.0f)) return 0u;
    uint32_t u = (uint32_t)q;
    return u > 262143u ? 262143u : u;
}

__device__ inline uint32_t seg_bound(uint64_t j, uint32_t n, uint32_t s) { return (uint32_t)((j * n) >> s); }

// first segment at depth s that contains position pos
__device__ inline uint32_t seg_of(uint32_t pos, uint32_t n, uint32_t s) {
    return (uint32_t)(((((uint64_t)pos + 1) << s) - 1) / n);
}

__device__ inline float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Bitonic passes with stride j < R of the merge sizes size_lo..size_hi on one
// thread's block of R consecutive keys (base a multiple of R): every such pair
// lies inside the block, so these passes run in registers with no barrier;
// the pairs and directions are those of the LDS passes.
template <uint32_t R>
__device__ inline void block_bitonic(uint32_t* keys, uint32_t base, uint32_t size_lo, uint32_t size_hi) {
    uint32_t k[R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) k[r] = keys[base + r];
    for (uint32_t size = size_lo; size <= size_hi; size <<= 1) {
#pragma unroll
        for (uint32_t j = R / 2; j > 0; j >>= 1) {
            if (j >= size) continue;
#pragma unroll
            for (uint32_t r = 0; r < R; ++r) {
                if (r & j) continue;
                const bool up = ((base + r) & size) == 0;
                const uint32_t lo = k[r], hi = k[r + j];
                const bool swap = (lo > hi) == up;
                k[r] = swap ? hi : lo;
                k[r + j] = swap ? lo : hi;
            }
        }
    }
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) keys[base + r] = k[r];
}

__device__ inline void block_bitonic_r(uint32_t R, uint32_t* keys, uint32_t tid, uint32_t size_lo, uint32_t size_hi) {
    switch (R) {
        case 2: block_bitonic<2>(keys, tid * 2, size_lo, size_hi); break;
        case 4: block_bitonic<4>(keys, tid * 4, size_lo, size_hi); break;
        case 8: block_bitonic<8>(keys, tid * 8, size_lo, size_hi); break;
        default: block_bitonic<16>(keys, tid * 16, size_lo, size_hi); break;
    }
}

__global__ __launch_bounds__(kBuildThreads) void build_slot_kernel(BuildArgs a) {
    __shared__ uint32_t lds[INSTA_MAX_TRIANGLES];   // 64 KiB: sort keys, then split scratch
    __shared__ uint16_t sL[INSTA_MAX_TRIANGLES];    // splits: one axis list (a copy of the global one)
    __shared__ uint16_t sLn[INSTA_MAX_TRIANGLES];   // splits: that list partitioned, copied out
    __shared__ float red[6][kBuildThreads / 64];
    const uint32_t tid = threadIdx.x;
    const uint32_t slot = a.first_slot + blockIdx.x;
    const uint32_t nf = a.nf;
    const float* V = a.verts + (size_t)blockIdx.x * a.nv * 3;
    float* VS = a.verts_store + (size_t)slot * a.vert_cap * 3;
    float4* cen = a.cen + (size_t)slot * a.tri_cap;
    float4* btris = a.btris + (size_t)slot * a.tri_cap * 3;
    float4* nodes = a.nodes + (size_t)slot * a.node_cap * 2;
    float4* proj = a.proj + (size_t)slot * a.tri_cap * 3;
    float4* proj_inv = a.proj_inv + (size_t)slot * a.tri_cap * 3;

    for (uint32_t i = tid; i < a.nv * 3; i += kBuildThreads) VS[i] = V[i];

    // 1. centroids and their bounds
    V3 lo = v3(INFINITY, INFINITY, INFINITY), hi = v3(-INFINITY, -INFINITY, -INFINITY);
    for (uint32_t i = tid; i < nf; i += kBuildThreads) {
        const int32_t* f = a.faces + 3 * i;
        V3 A = load_v(V, f[0]), B = load_v(V, f[1]), C = load_v(V, f[2]);
        V3 c = v3((A.x + B.x + C.x) / 3.0f, (A.y + B.y + C.y) / 3.0f, (A.z + B.z + C.z) / 3.0f);
        cen[i] = make_float4(c.x, c.y, c.z, 0.0f);
        lo = vmin(lo, c);
        hi = vmax(hi, c);
    }
    const uint32_t wave = tid / 64, lane = tid % 64;
    {
        float r[6] = {wave_min(lo.x), wave_min(lo.y), wave_min(lo.z), wave_max(hi.x), wave_max(hi.y), wave_max(hi.z)};
        if (lane == 0)
            for (int q = 0; q < 6; ++q) red[q][wave] = r[q];
    }
    __syncthreads();
    V3 cmin = v3(INFINITY, INFINITY, INFINITY), cmax = v3(-INFINITY, -INFINITY, -INFINITY);
    for (uint32_t w = 0; w < kBuildThreads / 64; ++w) {
        cmin = vmin(cmin, v3(red[0][w], red[1][w], red[2][w]));
        cmax = vmax(cmax, v3(red[3][w], red[4][w], red[5][w]));
    }
    const V3 ext = cmax - cmin;

    // 2. three presorted id lists (quantised centroid, id)
    uint32_t npad = 1;
    while (npad < nf) npad <<= 1;
    uint32_t* keys = reinterpret_cast<uint32_t*>(lds);
    uint16_t* lists = a.lists + (size_t)slot * 6 * a.tri_cap;     // [buffer][axis][tri_cap]
    uint32_t* qc = a.qcoord + (size_t)slot * 3 * a.tri_cap;
    for (uint32_t ax = 0; ax < 3; ++ax) {
        const float clo = ax == 0 ? cmin.x : (ax == 1 ? cmin.y : cmin.z);
        const float cex = ax == 0 ? ext.x : (ax == 1 ? ext.y : ext.z);
        for (uint32_t i = tid; i < npad; i += kBuildThreads) {
            uint32_t key = 0xffffffffu;
            if (i < nf) {
                const float4 c = cen[i];
                const uint32_t q = quantise18(ax == 0 ? c.x : (ax == 1 ? c.y : c.z), clo, cex);
                qc[ax * a.tri_cap + i] = q;
                key = (q << 14) | i;
            }
            keys[i] = key;
        }
        __syncthreads();
        // keys per thread for the register passes (1: every pass through LDS)
        const uint32_t R = npad >= 2 * kBuildThreads ? npad / kBuildThreads : 1u;
        uint32_t size0 = 2;
        if (R > 1) {
            block_bitonic_r(R, keys, tid, 2, R);  // merge sizes 2..R: all their passes
            __syncthreads();
            size0 = 2 * R;
        }
        for (uint32_t size = size0; size <= npad; size <<= 1) {
            for (uint32_t j = size >> 1; j >= R; j >>= 1) {
                for (uint32_t i = tid; i < npad / 2; i += kBuildThreads) {
                    const uint32_t x = ((i & ~(j - 1)) << 1) | (i & (j - 1));
                    const uint32_t y = x + j;
                    const bool up = (x & size) == 0;
                    const uint32_t kx = keys[x], ky = keys[y];
                    if ((kx > ky) == up) {
                        keys[x] = ky;
                        keys[y] = kx;
                    }
                }
                __syncthreads();
            }
            if (R > 1) {
                block_bitonic_r(R, keys, tid, size, size);  // strides R/2..1
                __syncthreads();
            }
        }
        for (uint32_t i = tid; i < nf; i += kBuildThreads) lists[ax * a.tri_cap + i] = (uint16_t)(keys[i] & 0x3fffu);
        __syncthreads();
    }

    // 3. median splits, one binary depth at a time.  LDS reuse: exclusive
    // left-counts (u16), right flags per face (u8), chosen axis per segment.
    uint16_t* excl = reinterpret_cast<uint16_t*>(lds);
    uint8_t* right = reinterpret_cast<uint8_t*>(lds) + 2 * INSTA_MAX_TRIANGLES;
    uint8_t* seg_axis = reinterpret_cast<uint8_t*>(lds) + 3 * INSTA_MAX_TRIANGLES;
    // positions per thread (a contiguous run: the scans below are blocked),
    // sized by the mesh so that every wave takes part
    const uint32_t per = (nf + kBuildThreads - 1) / kBuildThreads;
    const uint32_t p0 = tid * per;
    uint32_t cur = 0;
    for (uint32_t s = 0; s < 2 * a.k; ++s) {
        const uint16_t* L = lists + (size_t)cur * 3 * a.tri_cap;
        uint16_t* Ln = lists + (size_t)(cur ^ 1u) * 3 * a.tri_cap;
        const uint32_t n_seg = 1u << s;
        for (uint32_t j = tid; j < n_seg; j += kBuildThreads) {
            const uint32_t b = seg_bound(j, nf, s), e = seg_bound(j + 1, nf, s);
            uint32_t best_ax = 0, best_ext = 0;
            for (uint32_t ax = 0; ax < 3; ++ax) {
                const uint16_t* La = L + ax * a.tri_cap;
                const uint32_t ea = qc[ax * a.tri_cap + La[e - 1]] - qc[ax * a.tri_cap + La[b]];
                if (ax == 0 || ea > best_ext) {
                    best_ext = ea;
                    best_ax = ax;
                }
            }
            seg_axis[j] = (uint8_t)best_ax;
        }
        __syncthreads();
        // right flags: each position's id in its segment's chosen list; the
        // lists are read through an LDS copy (coalesced global loads)
        for (uint32_t ax = 0; ax < 3; ++ax) {
            const uint16_t* La = L + ax * a.tri_cap;
            for (uint32_t i = tid; i < nf; i += kBuildThreads) sL[i] = La[i];
            __syncthreads();
            if (p0 < nf) {
                uint32_t j = seg_of(p0, nf, s);
                uint32_t e = seg_bound(j + 1, nf, s);
                for (uint32_t pos = p0; pos < p0 + per && pos < nf; ++pos) {
                    while (pos >= e) {
                        ++j;
                        e = seg_bound(j + 1, nf, s);
                    }
                    if (seg_axis[j] == ax) right[sL[pos]] = pos >= seg_bound(2 * j + 1, nf, s + 1) ? 1 : 0;
                }
            }
            __syncthreads();
        }
        for (uint32_t ax = 0; ax < 3; ++ax) {
            const uint16_t* La = L + ax * a.tri_cap;
            uint16_t* Lna = Ln + ax * a.tri_cap;
            for (uint32_t i = tid; i < nf; i += kBuildThreads) sL[i] = La[i];
            __syncthreads();
            // block-wide exclusive scan of left flags in this list's order
            uint32_t cnt = 0;
            for (uint32_t pos = p0; pos < p0 + per && pos < nf; ++pos) cnt += right[sL[pos]] ? 0u : 1u;
            uint32_t incl = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t up = __shfl_up(incl, o);
                if (lane >= (uint32_t)o) incl += up;
            }
            if (lane == 63) red[0][wave] = __uint_as_float(incl);
            __syncthreads();
            uint32_t run = incl - cnt;
            for (uint32_t w = 0; w < wave; ++w) run += __float_as_uint(red[0][w]);
            for (uint32_t pos = p0; pos < p0 + per && pos < nf; ++pos) {
                excl[pos] = (uint16_t)run;
                run += right[sL[pos]] ? 0u : 1u;
            }
            __syncthreads();
            if (p0 < nf) {
                uint32_t j = seg_of(p0, nf, s);
                uint32_t e = seg_bound(j + 1, nf, s);
                for (uint32_t pos = p0; pos < p0 + per && pos < nf; ++pos) {
                    while (pos >= e) {
                        ++j;
                        e = seg_bound(j + 1, nf, s);
                    }
                    const uint32_t b = seg_bound(j, nf, s), mid = seg_bound(2 * j + 1, nf, s + 1);
                    const uint32_t id = sL[pos];
                    const uint32_t left_before = (uint32_t)excl[pos] - (uint32_t)excl[b];
                    const uint32_t np = right[id] ? mid + (pos - b) - left_before : b + left_before;
                    sLn[np] = (uint16_t)id;
                }
            }
            __syncthreads();
            for (uint32_t i = tid; i < nf; i += kBuildThreads) Lna[i] = sLn[i];
        }
        __syncthreads();  // the new lists, written by every thread, are read next
        cur ^= 1u;
    }
    const uint16_t* order = lists + (size_t)cur * 3 * a.tri_cap;  // axis-0 list: BVH order

    // 4. BVH-order triangle records
    for (uint32_t j = tid; j < nf; j += kBuildThreads) {
        const uint32_t o = order[j];
        const int32_t* f = a.faces + 3 * o;
        V3 A = load_v(V, f[0]), B = load_v(V, f[1]), C = load_v(V, f[2]);
        btris[3 * j + 0] = make_float4(A.x, A.y, A.z, B.x);
        btris[3 * j + 1] = make_float4(B.y, B.z, C.x, C.y);
        btris[3 * j + 2] = make_float4(C.z, __uint_as_float(o), __uint_as_float(a.masks[o]), 0.0f);
    }
    // leaf boxes
    const uint32_t n_leaves = 1u << (2 * a.k);
    for (uint32_t j = tid; j < n_leaves; j += kBuildThreads) {
        const uint32_t b = (uint32_t)(((uint64_t)j * nf) >> (2 * a.k));
        const uint32_t e = (uint32_t)(((uint64_t)(j + 1) * nf) >> (2 * a.k));
        V3 blo = v3(INFINITY, INFINITY, INFINITY), bhi = v3(-INFINITY, -INFINITY, -INFINITY);
        for (uint32_t t = b; t < e; ++t) {
            const int32_t* f = a.faces + 3 * (uint32_t)order[t];
            V3 A = load_v(V, f[0]), B = load_v(V, f[1]), C = load_v(V, f[2]);
            blo = vmin(blo, vmin(A, vmin(B, C)));
            bhi = vmax(bhi, vmax(A, vmax(B, C)));
        }
        const uint32_t node = a.leaf_base + j;
        nodes[2 * node] = make_float4(blo.x, blo.y, blo.z, __int_as_float((int32_t)b));
        nodes[2 * node + 1] = make_float4(bhi.x, bhi.y, bhi.z, __int_as_float((int32_t)e));
    }
    __syncthreads();
    // inner boxes, one heap level at a time
    for (int d = (int)a.k - 1; d >= 0; --d) {
        const uint32_t count = 1u << (2 * d);
        const uint32_t base = (count - 1) / 3;
        for (uint32_t p = tid; p < count; p += kBuildThreads) {
            const uint32_t node = base + p;
            V3 blo = v3(INFINITY, INFINITY, INFINITY), bhi = v3(-INFINITY, -INFINITY, -INFINITY);
            for (uint32_t c = 0; c < 4; ++c) {
                const uint32_t ch = 4 * node + 1 + c;
                blo = vmin(blo, xyz(nodes[2 * ch]));
                bhi = vmax(bhi, xyz(nodes[2 * ch + 1]));
            }
            nodes[2 * node] = make_float4(blo.x, blo.y, blo.z, __int_as_float(-1));
            nodes[2 * node + 1] = make_float4(bhi.x, bhi.y, bhi.z, __int_as_float(-1));
        }
        __syncthreads();
    }

    // 5. per-triangle frames in FP64 (original face order)
    for (uint32_t i = tid; i < nf; i += kBuildThreads) {
        const int32_t* f = a.faces + 3 * i;
        D3 A = {V[3 * f[0]], V[3 * f[0] + 1], V[3 * f[0] + 2]};
        D3 B = {V[3 * f[1]], V[3 * f[1] + 1], V[3 * f[1] + 2]};
        D3 C = {V[3 * f[2]], V[3 * f[2] + 1], V[3 * f[2] + 2]};
        double s = 1.0;
        if (slot != 0) {
            const float* CV = a.canon_verts;
            D3 cA = {CV[3 * f[0]], CV[3 * f[0] + 1], CV[3 * f[0] + 2]};
            D3 cB = {CV[3 * f[1]], CV[3 * f[1] + 1], CV[3 * f[1] + 2]};
            D3 cC = {CV[3 * f[2]], CV[3 * f[2] + 1], CV[3 * f[2] + 2]};
            D3 cr = dcross(dsub(B, A), dsub(C, A)), ccr = dcross(dsub(cB, cA), dsub(cC, cA));
            s = (0.5 * sqrt(ddot(cr, cr))) / (0.5 * sqrt(ddot(ccr, ccr)));
        }
        // Triangle::tbn (triangle.cuh:43-57): N = |(b-a) x (c-a)|, B = |(b-a) x N|, T = |B x N|
        D3 t = dsub(B, A);
        D3 n = dnormalized(dcross(t, dsub(C, A)));
        D3 bi = dnormalized(dcross(t, n));
        D3 ta = dnormalized(dcross(bi, n));
        const double sc = (a.masks[i] == 1u) ? s : 1.0;
        double m[12] = {ta.x * sc, bi.x * sc, n.x * sc, A.x, ta.y * sc, bi.y * sc, n.y * sc, A.y,
                        ta.z * sc, bi.z * sc, n.z * sc, A.z};
        // affine inverse by cofactors
        double c00 = m[5] * m[10] - m[6] * m[9], c01 = m[2] * m[9] - m[1] * m[10], c02 = m[1] * m[6] - m[2] * m[5];
        double c10 = m[6] * m[8] - m[4] * m[10], c11 = m[0] * m[10] - m[2] * m[8], c12 = m[2] * m[4] - m[0] * m[6];
        double c20 = m[4] * m[9] - m[5] * m[8], c21 = m[1] * m[8] - m[0] * m[9], c22 = m[0] * m[5] - m[1] * m[4];
        double det = m[0] * c00 + m[1] * c10 + m[2] * c20;
        double id = 1.0 / det;
        double inv[9] = {c00 * id, c01 * id, c02 * id, c10 * id, c11 * id, c12 * id, c20 * id, c21 * id, c22 * id};
        double it[3];
        for (int r = 0; r < 3; ++r) it[r] = -(inv[3 * r] * m[3] + inv[3 * r + 1] * m[7] + inv[3 * r + 2] * m[11]);
        for (int r = 0; r < 3; ++r) {
            proj[3 * i + r] = make_float4((float)m[4 * r], (float)m[4 * r + 1], (float)m[4 * r + 2], (float)m[4 * r + 3]);
            proj_inv[3 * i + r] = make_float4((float)inv[3 * r], (float)inv[3 * r + 1], (float)inv[3 * r + 2], (float)it[r]);
        }
    }
}

__global__ void blend_expression_kernel(const float* __restrict__ canon, const float* __restrict__ bases,
                                        const float* __restrict__ coeffs, uint32_t n_bases, uint32_t n, float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = canon[i];
    for (uint32_t k = 0; k < n_bases; ++k) acc = acc + coeffs[k] * bases[(size_t)k * n + i];
    out[i] = acc;
}

constexpr uint32_t kQueryThreads = 256;

// nearest triangle per point: the frame's candidate lists where they cover
// the point (use_lists), else the BVH.  kCount: also sum the query work into
// counts[kTravCounters] (a separate build for measurement runs)
template <bool kCount>
__global__ __launch_bounds__(kQueryThreads) void closest_kernel(MeshView m, CellListView cl, const float* __restrict__ pts,
                                                               uint32_t n, int32_t* __restrict__ ids,
                                                               float* __restrict__ dist,
                                                               unsigned long long* __restrict__ counts) {
    extern __shared__ uint32_t stack[];  // [stack_depth(k)][kQueryThreads]
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    TraversalCount tc;
    if (i < n) {
        V3 p = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        Hit h = closest_impl<kCount>(m, cl, p, stack + threadIdx.x, kQueryThreads, &tc);
        ids[i] = h.idx < 0 ? -1 : (int32_t)h.orig;
        if (dist) dist[i] = sqrtf(h.dist_sq);
    }
    if constexpr (kCount) {
        unsigned long long v[kTravCounters] = {i < n ? 1ull : 0ull, tc.inner, tc.leaves, tc.triangles, tc.lists,
                                               tc.entries};
#pragma unroll
        for (int c = 0; c < kTravCounters; ++c) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
            if ((threadIdx.x & 63u) == 0 && v[c]) atomicAdd(counts + c, v[c]);
        }
    }
}

// mark_flame_density_grid + grid_to_bitfield (testbed_nerf.cu:364-389, 585-609):
// one thread per byte of 8 Morton-consecutive cells.
__global__ __launch_bounds__(kQueryThreads) void occupancy_kernel(MeshView m, uint32_t G, float threshold,
                                                                 uint8_t* __restrict__ bits, uint32_t n_cells) {
    extern __shared__ uint32_t stack[];  // [stack_depth(k)][kQueryThreads]
    // one cell per lane, cells in Morton order; the wave's 64 cell bits are
    // gathered with a ballot and stored as 8 bytes
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const float g = (float)G;
    // cells farther than the threshold are never occupied: bound the query
    // a little above threshold^2 so sqrt(bound) >= threshold after rounding
    const float bound = threshold * threshold * (1.0f + 0x1p-10f);
    bool occ = false;
    if (idx < n_cells) {
        const uint32_t x = compact3(idx), y = compact3(idx >> 1), z = compact3(idx >> 2);
        V3 pos = v3((((float)x + 0.5f) / g - 0.5f) * 1.0f + 0.5f, (((float)y + 0.5f) / g - 0.5f) * 1.0f + 0.5f,
                    (((float)z + 0.5f) / g - 0.5f) * 1.0f + 0.5f);
        Hit h = bvh_closest(m, pos, stack + threadIdx.x, kQueryThreads, bound);
        if (h.idx >= 0) {
            const float dist = sqrtf(h.dist_sq);
            const uint32_t mask = __float_as_uint(m.btris[3 * h.idx + 2].z);
            occ = dist < threshold && mask != 0u;
        }
    }
    const uint64_t word = __ballot(occ);
    const uint32_t lane = threadIdx.x & 63u;
    if ((lane & 7u) == 0 && idx < n_cells) bits[idx >> 3] = (uint8_t)(word >> lane);
}

// Linear-order copy of a Morton-order bitfield: bit x + G(y + G z) of the
// output is bit morton3(x, y, z) of the input.  One thread per output byte.
__global__ void grid_linear_kernel(const uint8_t* __restrict__ morton_bits, uint32_t G, uint8_t* __restrict__ lin,
                                   uint32_t n_bytes) {
    const uint32_t byte = blockIdx.x * blockDim.x + threadIdx.x;
    if (byte >= n_bytes) return;
    uint32_t v = 0;
    for (uint32_t b = 0; b < 8; ++b) {
        const uint32_t cell = byte * 8 + b;
        const uint32_t x = cell % G, y = (cell / G) % G, z = cell / (G * G);
        const uint32_t m = morton3(x, y, z);
        v |= ((morton_bits[m >> 3] >> (m & 7u)) & 1u) << b;
    }
    lin[byte] = (uint8_t)v;
}

}  // namespace

insta_status launch_grid_linear(insta_ctx* ctx, uint32_t slot, hipStream_t stream) {
    const uint32_t G = ctx->cfg.grid_size;
    const uint32_t n_bytes = G * G * G / 8;
    hipLaunchKernelGGL(grid_linear_kernel, dim3((n_bytes + 255) / 256), dim3(256), 0, stream, ctx->grid(slot), G,
                       ctx->grid_lin(slot), n_bytes);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_build_slots(insta_ctx* ctx, uint32_t first_slot, uint32_t n_slots, const float* verts,
                                hipStream_t stream) {
    BuildArgs a;
    a.verts = verts;
    a.canon_verts = ctx->d_verts;
    a.faces = ctx->d_faces;
    a.masks = ctx->d_masks;
    a.verts_store = ctx->d_verts;
    a.nodes = ctx->d_nodes;
    a.btris = ctx->d_btris;
    a.proj = ctx->d_proj;
    a.proj_inv = ctx->d_proj_inv;
    a.cen = ctx->d_cen;
    a.lists = ctx->d_build_lists;
    a.qcoord = ctx->d_build_q;
    a.first_slot = first_slot;
    a.nv = ctx->nv;
    a.nf = ctx->nf;
    a.k = ctx->bvh_k;
    a.leaf_base = ctx->leaf_base;
    a.node_cap = ctx->node_cap;
    a.tri_cap = ctx->cfg.max_triangles;
    a.vert_cap = ctx->cfg.max_vertices;
    if (ctx->nf == 0 || ctx->nf > INSTA_MAX_TRIANGLES || first_slot + n_slots > ctx->n_slots) {
        set_error("build: mesh size or slot range out of bounds");
        return INSTA_ERR_INVALID;
    }
    // the slots' candidate lists describe the old meshes until rebuilt
    insta_status st = launch_cell_lists_invalidate(ctx, first_slot, n_slots, stream);
    if (st != INSTA_OK) return st;
    hipLaunchKernelGGL(build_slot_kernel, dim3(n_slots), dim3(kBuildThreads), 0, stream, a);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_blend_expression(insta_ctx* ctx, uint32_t slot, const float* bases, uint32_t n_bases,
                                     const float* coeffs, hipStream_t stream) {
    // blended straight into the slot's vertex store (the build's vertex copy
    // is then in place), so concurrent expressions on other slots never share
    // a buffer
    const uint32_t n = ctx->nv * 3;
    float* dst = ctx->d_verts + (size_t)slot * ctx->cfg.max_vertices * 3;
    hipLaunchKernelGGL(blend_expression_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, ctx->d_verts, bases,
                       coeffs, n_bases, n, dst);
    INSTA_HIP(hipGetLastError());
    return launch_build_slots(ctx, slot, 1, dst, stream);
}

insta_status launch_closest(insta_ctx* ctx, uint32_t slot, const float* pts, uint32_t n, int32_t* ids, float* dist,
                            bool use_lists, hipStream_t stream) {
    if (n == 0) return INSTA_OK;
    const size_t shmem = (size_t)stack_depth(ctx->bvh_k) * kQueryThreads * sizeof(uint32_t);
    const CellListView cl = use_lists ? ctx->cells(slot) : CellListView{};
    const dim3 grid((n + kQueryThreads - 1) / kQueryThreads);
    if (ctx->d_trav_counts)
        hipLaunchKernelGGL(closest_kernel<true>, grid, dim3(kQueryThreads), shmem, stream, ctx->mesh(slot), cl, pts, n,
                           ids, dist, ctx->d_trav_counts);
    else
        hipLaunchKernelGGL(closest_kernel<false>, grid, dim3(kQueryThreads), shmem, stream, ctx->mesh(slot), cl, pts, n,
                           ids, dist, nullptr);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_occupancy(insta_ctx* ctx, uint32_t slot, float threshold, hipStream_t stream) {
    const uint32_t G = ctx->cfg.grid_size;
    const uint32_t n_cells = G * G * G;  // G is a multiple of 8: whole bytes per 8 lanes
    const size_t shmem = (size_t)stack_depth(ctx->bvh_k) * kQueryThreads * sizeof(uint32_t);
    hipLaunchKernelGGL(occupancy_kernel, dim3((n_cells + kQueryThreads - 1) / kQueryThreads), dim3(kQueryThreads),
                       shmem, stream, ctx->mesh(slot), G, threshold, ctx->grid(slot), n_cells);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

}  // namespace insta
