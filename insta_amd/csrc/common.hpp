// common.hpp -- device math shared by the INSTA hot-path kernels (gfx950).
//
// All sources are compiled with -ffp-contract=off: each float operation below
// rounds once, in the order written.  The geometric predicates (triangle and
// box distances, Morton codes, march stepping) must give bit-identical results
// on identical inputs, so their operation order is part of the contract.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/insta_capi.h"

#define INSTA_DEV __device__ __forceinline__
#define INSTA_HD __host__ __device__ __forceinline__

namespace insta {

struct V3 {
    float x, y, z;
};
INSTA_HD V3 v3(float x, float y, float z) { return {x, y, z}; }
INSTA_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
INSTA_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
INSTA_HD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
INSTA_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
INSTA_HD V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
INSTA_HD float sqnorm(V3 a) { return dot(a, a); }
INSTA_DEV float norm(V3 a) { return sqrtf(sqnorm(a)); }
INSTA_DEV V3 vmin(V3 a, V3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
INSTA_DEV V3 vmax(V3 a, V3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
INSTA_DEV float signf(float x) { return copysignf(1.0f, x); }
INSTA_DEV float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
INSTA_DEV V3 xyz(float4 a) { return {a.x, a.y, a.z}; }

// Point-triangle squared distance (Inigo Quilez's formulation, as in
// instant-ngp/include/neural-graphics-primitives/triangle.cuh:88-107).
INSTA_DEV float tri_distance_sq(V3 a, V3 b, V3 c, V3 pos) {
    V3 v21 = b - a; V3 p1 = pos - a;
    V3 v32 = c - b; V3 p2 = pos - b;
    V3 v13 = a - c; V3 p3 = pos - c;
    V3 nor = cross(v21, v13);
    float s = signf(dot(cross(v21, nor), p1)) + signf(dot(cross(v32, nor), p2)) + signf(dot(cross(v13, nor), p3));
    if (s < 2.0f) {
        float e1 = sqnorm(v21 * clampf(dot(v21, p1) / sqnorm(v21), 0.0f, 1.0f) - p1);
        float e2 = sqnorm(v32 * clampf(dot(v32, p2) / sqnorm(v32), 0.0f, 1.0f) - p2);
        float e3 = sqnorm(v13 * clampf(dot(v13, p3) / sqnorm(v13), 0.0f, 1.0f) - p3);
        return fminf(fminf(e1, e2), e3);
    }
    float d = dot(nor, p1);
    return d * d / sqnorm(nor);
}

// Point-AABB squared distance (bounding_box.cuh:228-230).
INSTA_DEV float box_distance_sq(float4 lo, float4 hi, V3 p) {
    float dx = fmaxf(fmaxf(lo.x - p.x, p.x - hi.x), 0.0f);
    float dy = fmaxf(fmaxf(lo.y - p.y, p.y - hi.y), 0.0f);
    float dz = fmaxf(fmaxf(lo.z - p.z, p.z - hi.z), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

// 10-bit-per-axis Morton interleave: bit i of x/y/z -> bit 3i/3i+1/3i+2.
INSTA_HD uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
INSTA_HD uint32_t compact3(uint32_t v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030c30c3u;
    v = (v | (v >> 4)) & 0x0300f00fu;
    v = (v | (v >> 8)) & 0x030000ffu;
    v = (v | (v >> 16)) & 0x000003ffu;
    return v;
}
INSTA_HD uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) { return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2); }

// PCG32 (O'Neill 2014, the pcg32 generator of Jakob's pcg32.h that
// instant-ngp uses as default_rng_t): 64-bit LCG state, XSH-RR output,
// O(log n) jump-ahead.  Host and device produce the same stream.
struct Pcg32 {
    static constexpr uint64_t kMult = 0x5851f42d4c957f2dull;
    static constexpr uint64_t kDefaultState = 0x853c49e6748fea9bull;
    static constexpr uint64_t kDefaultStream = 0xda3e39cb94b95bdbull;
    uint64_t state, inc;

    INSTA_HD static Pcg32 seeded(uint64_t initstate, uint64_t initseq = kDefaultStream) {
        Pcg32 r{0u, (initseq << 1u) | 1u};
        r.next_uint();
        r.state += initstate;
        r.next_uint();
        return r;
    }
    INSTA_HD uint32_t next_uint() {
        const uint64_t old = state;
        state = old * kMult + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    // uniform in [0, 1): 23 random mantissa bits
    INSTA_HD float next_float() {
        const uint32_t u = (next_uint() >> 9) | 0x3f800000u;
        float f;
        __builtin_memcpy(&f, &u, 4);
        return f - 1.0f;
    }
    INSTA_HD void advance(uint64_t delta) {
        uint64_t cur_mult = kMult, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
        while (delta > 0) {
            if (delta & 1u) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1u) * cur_plus;
            cur_mult *= cur_mult;
            delta /= 2u;
        }
        state = acc_mult * state + acc_plus;
    }
};

// ---------------------------------------------------------------------------
// Mesh storage.  Slot 0 is the canonical mesh, slot f+1 posed frame f.  All
// slots share the topology, so the BVH shape (levels, node count) is common.
//   nodes  [slot][n_nodes][2] float4: (min xyz, begin) (max xyz, end); begin/
//          end are int bit patterns of the BVH-order triangle range (leaves)
//   btris  [slot][nf][3] float4: (a xyz, b.x) (b.yz, c.xy) (c.z, orig id,
//          mask id, 0) in BVH order; id/mask are uint bit patterns
//   proj / proj_inv [slot][nf][3] float4: rows of the 3x4 frame [A | a]
//   cen    [slot][nf] float4: float centroid xyz (original face order)
// ---------------------------------------------------------------------------
struct MeshView {
    const float4* nodes;
    const float4* btris;
    const float4* proj;
    const float4* proj_inv;
    const float4* cen;
    uint32_t nf, leaf_base, k;
};

struct Hit {
    int32_t idx;       // BVH-order index, -1 if none
    uint32_t orig;     // original face id
    float dist_sq;
};

// Nearest triangle (bvh.cuh:48-96), exact under the (dist_sq, original id)
// order: the result equals the brute-force argmin, whatever the visiting
// order, because a subtree is skipped only when its box lies beyond the best
// triangle found so far by more than a rounding margin.
//
// Traversal (implicit 4-ary heap): descend into the nearest surviving child,
// pushing the other survivors far-to-near; at a leaf rank its triangles, then
// pop until an entry whose box distance is still within the (possibly
// tightened) bound.  Inner-node descents and leaf tests run as two loops so
// the lanes of a wave execute the same code path together.
//
// `stack` points at this thread's column of a [stack_depth(k)][stride] uint32
// LDS array.  An entry packs the child's box distance, truncated to its upper
// 16 bits (never above the true distance, so the pop-time check stays
// conservative), with the 16-bit node index.
constexpr uint32_t kNoNode = 0xffffffffu;

// conservative prune bound (relative 2^-18 + absolute 2^-26): rounding can
// put a box a few ulp beyond the best triangle it contains
INSTA_DEV float prune_bound(float best) { return best + best * 0x1p-18f + 0x1p-26f; }

INSTA_DEV uint32_t pop_live(const uint32_t* stack, int& sp, uint32_t stride, float bound) {
    while (sp > 0) {
        --sp;
        const uint32_t w = stack[sp * stride];
        if (__uint_as_float(w & 0xffff0000u) <= bound) return w & 0xffffu;
    }
    return kNoNode;
}

// traversal work of one query (counting builds only)
struct TraversalCount {
    uint32_t inner = 0, leaves = 0, triangles = 0, lists = 0, entries = 0;
};
// device counters of a counting build: queries, inner nodes, leaves, exact
// triangle distances, list queries, list entries scanned
constexpr int kTravCounters = 6;

template <bool kCount>
INSTA_DEV Hit bvh_closest_impl(const MeshView& m, V3 p, uint32_t* stack, uint32_t stride, float max_dist_sq,
                               TraversalCount* tc) {
    // max_dist_sq (bvh.cuh:48): only triangles closer than this are reported
    float best = max_dist_sq;
    int32_t best_idx = -1;
    uint32_t best_orig = 0xffffffffu;
    float bound = prune_bound(best);
    const uint32_t shift = 2u * m.k;
    int sp = 0;
    uint32_t node = 0;
    for (;;) {
        while (node < m.leaf_base) {
            if constexpr (kCount) ++tc->inner;
            const uint32_t first = 4u * node + 1u;
            float cd[4];
            uint32_t ci[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                ci[c] = first + c;
                cd[c] = box_distance_sq(m.nodes[2 * (first + c)], m.nodes[2 * (first + c) + 1], p);
            }
            // sorting network, farthest first (bvh.cuh:8-46)
#define INSTA_CAS(i, j)                              \
    if (cd[i] < cd[j]) {                             \
        float td = cd[i]; cd[i] = cd[j]; cd[j] = td; \
        uint32_t ti = ci[i]; ci[i] = ci[j]; ci[j] = ti; \
    }
            INSTA_CAS(0, 2) INSTA_CAS(1, 3) INSTA_CAS(0, 1) INSTA_CAS(2, 3) INSTA_CAS(1, 2)
#undef INSTA_CAS
            uint32_t next = kNoNode;
            float nd = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if (cd[c] <= bound) {
                    if (next != kNoNode) {
                        stack[sp * stride] = (__float_as_uint(nd) & 0xffff0000u) | next;
                        ++sp;
                    }
                    next = ci[c];
                    nd = cd[c];
                }
            }
            node = next != kNoNode ? next : pop_live(stack, sp, stride, bound);
        }
        if (node == kNoNode) break;
        const uint32_t j = node - m.leaf_base;
        const uint32_t b = (uint32_t)(((uint64_t)j * m.nf) >> shift);
        const uint32_t e = (uint32_t)(((uint64_t)(j + 1) * m.nf) >> shift);
        if constexpr (kCount) ++tc->leaves;
        for (uint32_t t = b; t < e; ++t) {
            const float4 r0 = m.btris[3 * t + 0];
            const float4 r1 = m.btris[3 * t + 1];
            const float4 r2 = m.btris[3 * t + 2];
            const V3 va = v3(r0.x, r0.y, r0.z), vb = v3(r0.w, r1.x, r1.y), vc = v3(r1.z, r1.w, r2.x);
            const float4 lo = make_float4(fminf(fminf(va.x, vb.x), vc.x), fminf(fminf(va.y, vb.y), vc.y),
                                          fminf(fminf(va.z, vb.z), vc.z), 0.0f);
            const float4 hi = make_float4(fmaxf(fmaxf(va.x, vb.x), vc.x), fmaxf(fmaxf(va.y, vb.y), vc.y),
                                          fmaxf(fmaxf(va.z, vb.z), vc.z), 0.0f);
            // the triangle's own box first: cheap, and conservative like the
            // node test (bound carries the same rounding margin)
            if (box_distance_sq(lo, hi, p) > bound) continue;
            if constexpr (kCount) ++tc->triangles;
            const float d = tri_distance_sq(va, vb, vc, p);
            const uint32_t o = __float_as_uint(r2.y);
            if (d < best || (d == best && o < best_orig)) {
                best = d;
                best_idx = (int32_t)t;
                best_orig = o;
                bound = prune_bound(best);
            }
        }
        bound = prune_bound(best);
        node = pop_live(stack, sp, stride, bound);
        if (node == kNoNode) break;
    }
    Hit h;
    h.idx = best_idx;
    h.orig = best_orig;
    h.dist_sq = best_idx < 0 ? 0.0f : best;
    return h;
}

INSTA_DEV Hit bvh_closest(const MeshView& m, V3 p, uint32_t* stack, uint32_t stride,
                          float max_dist_sq = 3.40282347e38f) {
    return bvh_closest_impl<false>(m, p, stack, stride, max_dist_sq, nullptr);
}

// ---------------------------------------------------------------------------
// Per-cell candidate lists.  For a frame, each occupied cell of its mesh
// occupancy grid (G^3, linear order x + G(y + G z)) is split into S^3
// sub-cells.  A sub-cell's list holds every triangle whose box lies within U
// of the sub-cell's box (grown by kCellEps), where U^2 bounds the squared
// distance from any point of the box to the triangle nearest to the box
// centre (the largest over the 8 corners: squared distance to a convex set
// is convex), with a relative margin.  A triangle left out is therefore
// farther from every point of the sub-cell than that triangle, and the
// nearest triangle of a point in the sub-cell is always on its list.
//   bits   [G^3/64] 64-bit words of the occupancy bits the lists were built on
//   rank   [G^3/64] occupied cells before each word
//   head   [cells * S^3] (first entry, count | seed << 16); kNoList: no list
//   ent    entries: (box distance^2 truncated to its upper 16 bits) << 16 |
//          triangle (BVH order); the seed (nearest at the centre) first, the
//          rest ascending
//   state  [kClState] occupied cells, entries used, valid (1: lists answer
//          queries), occupied cells beyond the cap, sub-cells without a list
// ---------------------------------------------------------------------------
constexpr uint32_t kNoList = 0xffffffffu;
constexpr uint32_t kClState = 8;
constexpr float kCellEps = 1e-5f;

struct CellListView {
    const uint64_t* bits;
    const uint32_t* rank;
    const uint2* head;
    const uint32_t* ent;
    const uint32_t* state;
    uint32_t G, S, cap;
};

// Sub-cell list of a point in [0,1]^3: false when the lists are not valid,
// the point's cell is not occupied, or its sub-cell overflowed.
INSTA_DEV bool cell_list_find(const CellListView& cl, V3 p, uint2& hd) {
    if (!cl.state || cl.state[2] == 0u) return false;
    // the cell the march's occupancy lookup gives the point (march.hpp), so
    // every marched sample finds its cell; the point lies in that cell's
    // box up to rounding far below kCellEps
    const float g = (float)cl.G;
    const float gx = ((p.x - 0.5f) * 1.0f + 0.5f) * g, gy = ((p.y - 0.5f) * 1.0f + 0.5f) * g,
                gz = ((p.z - 0.5f) * 1.0f + 0.5f) * g;
    const uint32_t hi = cl.G - 1u;
    const uint32_t cx = min((uint32_t)gx, hi), cy = min((uint32_t)gy, hi), cz = min((uint32_t)gz, hi);
    const uint32_t cell = cx + cl.G * (cy + cl.G * cz);
    const uint64_t w = cl.bits[cell >> 6];
    const uint32_t b = cell & 63u;
    if (!((w >> b) & 1ull)) return false;
    const uint32_t r = cl.rank[cell >> 6] + (uint32_t)__popcll(w & ((1ull << b) - 1ull));
    if (r >= cl.cap) return false;
    const float s = (float)cl.S;
    const uint32_t sm = cl.S - 1u;
    const uint32_t sx = min((uint32_t)((gx - (float)cx) * s), sm);
    const uint32_t sy = min((uint32_t)((gy - (float)cy) * s), sm);
    const uint32_t sz = min((uint32_t)((gz - (float)cz) * s), sm);
    hd = cl.head[(size_t)r * (cl.S * cl.S * cl.S) + sx + cl.S * (sy + cl.S * sz)];
    return hd.y != kNoList;
}

// Nearest triangle among a sub-cell's list, exact under the same
// (dist_sq, original id) order and prune margin as bvh_closest_impl: the scan
// stops at the first entry whose box distance exceeds the bound (entries
// ascend), and a triangle is measured only when both its own box and its
// plane lie within the bound (the distance to a triangle is at least the
// distance to its plane).  The scan is latency-bound (each entry is a load of
// the entry, then of its triangle), so entries go in batches of kListBatch:
// the batch's triangles are loaded together, with the next batch's entry
// words already in flight, then tested in order; loads past the stopping
// entry are wasted, never used.
constexpr uint32_t kListBatch = 3;

template <bool kCount>
INSTA_DEV Hit list_closest_impl(const MeshView& m, const uint32_t* __restrict__ ent, uint2 hd, V3 p,
                                TraversalCount* tc) {
    float best = 3.40282347e38f;
    int32_t best_idx = -1;
    uint32_t best_orig = 0xffffffffu;
    float bound = prune_bound(best);
    const uint32_t n = hd.y & 0xffffu;
    const uint32_t* e = ent + hd.x;
    if constexpr (kCount) ++tc->lists;
    bool done = false;
    uint32_t wn[kListBatch];
#pragma unroll
    for (uint32_t j = 0; j < kListBatch; ++j) wn[j] = e[min(j, n - 1u)];
    for (uint32_t k = 0; k < n && !done; k += kListBatch) {
        uint32_t w[kListBatch];
        float4 R[kListBatch][3];
#pragma unroll
        for (uint32_t j = 0; j < kListBatch; ++j) w[j] = wn[j];
        // the next batch's entry words, in flight with this batch's triangles
#pragma unroll
        for (uint32_t j = 0; j < kListBatch; ++j) wn[j] = e[min(k + kListBatch + j, n - 1u)];
#pragma unroll
        for (uint32_t j = 0; j < kListBatch; ++j) {
            const uint32_t t = w[j] & 0xffffu;
            R[j][0] = m.btris[3 * t + 0];
            R[j][1] = m.btris[3 * t + 1];
            R[j][2] = m.btris[3 * t + 2];
        }
#pragma unroll
        for (uint32_t j = 0; j < kListBatch; ++j) {
            if (done || k + j >= n || __uint_as_float(w[j] & 0xffff0000u) > bound) {
                done = true;
                continue;
            }
            if constexpr (kCount) ++tc->entries;
            const float4 r0 = R[j][0], r1 = R[j][1], r2 = R[j][2];
            const V3 va = v3(r0.x, r0.y, r0.z), vb = v3(r0.w, r1.x, r1.y), vc = v3(r1.z, r1.w, r2.x);
            const float4 lo = make_float4(fminf(fminf(va.x, vb.x), vc.x), fminf(fminf(va.y, vb.y), vc.y),
                                          fminf(fminf(va.z, vb.z), vc.z), 0.0f);
            const float4 hi = make_float4(fmaxf(fmaxf(va.x, vb.x), vc.x), fmaxf(fmaxf(va.y, vb.y), vc.y),
                                          fmaxf(fmaxf(va.z, vb.z), vc.z), 0.0f);
            if (box_distance_sq(lo, hi, p) > bound) continue;
            const V3 nor = cross(vb - va, va - vc);
            const float pd = dot(nor, p - va);
            if (pd * pd > bound * sqnorm(nor) * (1.0f + 0x1p-8f)) continue;
            if constexpr (kCount) ++tc->triangles;
            const float d = tri_distance_sq(va, vb, vc, p);
            const uint32_t o = __float_as_uint(r2.y);
            if (d < best || (d == best && o < best_orig)) {
                best = d;
                best_idx = (int32_t)(w[j] & 0xffffu);
                best_orig = o;
                bound = prune_bound(best);
            }
        }
    }
    Hit h;
    h.idx = best_idx;
    h.orig = best_orig;
    h.dist_sq = best_idx < 0 ? 0.0f : best;
    return h;
}

// Nearest triangle of a point: its sub-cell's list when there is one, else
// the BVH (same result).
template <bool kCount>
INSTA_DEV Hit closest_impl(const MeshView& m, const CellListView& cl, V3 p, uint32_t* stack, uint32_t stride,
                           TraversalCount* tc) {
    uint2 hd;
    const bool in_box = p.x >= 0.0f && p.x <= 1.0f && p.y >= 0.0f && p.y <= 1.0f && p.z >= 0.0f && p.z <= 1.0f;
    if (in_box && cell_list_find(cl, p, hd)) return list_closest_impl<kCount>(m, cl.ent, hd, p, tc);
    return bvh_closest_impl<kCount>(m, p, stack, stride, 3.40282347e38f, tc);
}

INSTA_DEV Hit closest(const MeshView& m, const CellListView& cl, V3 p, uint32_t* stack, uint32_t stride) {
    return closest_impl<false>(m, cl, p, stack, stride, nullptr);
}

// Stack entries a query can hold: at most 3 pending siblings per inner level
// of a tree with k inner levels (k <= 6 for 16384 triangles).  Kernels keep
// the stack in dynamic LDS sized by this, so shallower trees leave room for
// more resident waves.
INSTA_HD uint32_t stack_depth(uint32_t k) { return k ? 3u * k : 1u; }

}  // namespace insta
