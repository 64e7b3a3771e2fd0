// oracle.cpp -- scalar CPU restatement of INSTA's per-sample hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the reported
// CPU baseline: tests/, the package's smoke() check and bench.py's cpu_baseline
// leg load it; nothing in the product path (insta_amd/) may call it.  It is
// compiled with -O3 -fopenmp -ffp-contract=off (no -ffast-math) so that every
// float operation rounds exactly once, in the order written, like the HIP
// kernels, which are compiled the same way.  Integer/index results (nearest
// triangle ids, hash-table indices, occupancy bits, march step counts) must
// then match the GPU bit for bit on identical inputs.
//
// Each function names the reference code it restates (file:line under the
// reference tree).  Where the reference calls into tiny-cuda-nn (an empty
// submodule in the reference tree: hash grid, MLPs, spherical harmonics), the
// published Instant-NGP algorithm (Mueller et al. 2022) is restated and the
// declared convention is written next to the code.
//
// Parity pinning (what checks each part; tests/test_oracle_*.py):
//   distance_sq / closest     brute force over all triangles, closed forms
//   bvh build                 invariants (cover, nesting, leaf ranges)
//   frames / inverse          orthonormality, M * M^-1 = I, rigid motion
//   warp                      identity pose => identity map; rigid pose =>
//                             exact rigid inverse; FP64 cross-check
//   occupancy                 brute-force distances, Morton round trip
//   ld_random_val / sobol     integer KATs from the recurrence and bit tricks
//   march                     empty grid => 0 steps; full grid => closed form
//   composite                 constant sigma closed form
//   hash encode               vertex => table entry; dense bijection; weights
//   MLP / SH                  one-hot closed forms; SH orthonormality
#include "../include/insta_capi.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

// Planted mistakes, for showing that the pinning tests are live
// (tests/test_oracle_pinning.py builds this file with -DORC_MUTANT=k into a
// scratch directory and expects each pinning check to fail on its mutant):
//   1  interpolate_transformation: self and neighbour factors swapped (1 <-> 4)
//   2  spherical harmonics: sign of the first band-1 coefficient flipped
//   3  Huber loss: threshold 0.1 -> 0.2
//   4  march: the one-step start jitter of advance_pos_nerf dropped
//   5  training target: background blended in stored (sRGB) values, not linear
//   6  render buffer: spp accumulation weights 1/spp instead of 1/(spp + 1)
//   7  depth supervision: target depth not scaled by the unnormalised ray direction's length
//   8  depth supervision: the depth suffix left out of the density gradient
// The default build (ORC_MUTANT 0) has none of them.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

namespace {

// ---------------------------------------------------------------------------
// small vector math (float, evaluation order fixed)
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return {x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float sqnorm(V3 a) { return dot(a, a); }
inline float norm(V3 a) { return std::sqrt(sqnorm(a)); }
inline V3 vmin(V3 a, V3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline float signf(float x) { return std::copysign(1.0f, x); }  // common.h:191-193
inline float clampf(float v, float lo, float hi) { return std::fmin(std::fmax(v, lo), hi); }

struct D3 {
    double x, y, z;
};
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 crossd(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Eigen::normalized(): unchanged when the norm is zero.
inline D3 normalized(D3 a) {
    double s = dotd(a, a);
    if (s > 0.0) {
        double n = std::sqrt(s);
        return {a.x / n, a.y / n, a.z / n};
    }
    return a;
}

// ---------------------------------------------------------------------------
// triangle distance (triangle.cuh:88-107) and box distance (bounding_box.cuh:228-230)
// ---------------------------------------------------------------------------
struct Tri {
    V3 a, b, c;
};

inline float tri_distance_sq(const Tri& t, V3 pos) {
    V3 v21 = t.b - t.a; V3 p1 = pos - t.a;
    V3 v32 = t.c - t.b; V3 p2 = pos - t.b;
    V3 v13 = t.a - t.c; V3 p3 = pos - t.c;
    V3 nor = cross(v21, v13);
    float s = signf(dot(cross(v21, nor), p1)) + signf(dot(cross(v32, nor), p2)) +
              signf(dot(cross(v13, nor), p3));
    if (s < 2.0f) {
        float e1 = sqnorm(v21 * clampf(dot(v21, p1) / sqnorm(v21), 0.0f, 1.0f) - p1);
        float e2 = sqnorm(v32 * clampf(dot(v32, p2) / sqnorm(v32), 0.0f, 1.0f) - p2);
        float e3 = sqnorm(v13 * clampf(dot(v13, p3) / sqnorm(v13), 0.0f, 1.0f) - p3);
        return std::fmin(std::fmin(e1, e2), e3);
    }
    float d = dot(nor, p1);
    return d * d / sqnorm(nor);
}

inline float box_distance_sq(V3 lo, V3 hi, V3 p) {
    float dx = std::fmax(std::fmax(lo.x - p.x, p.x - hi.x), 0.0f);
    float dy = std::fmax(std::fmax(lo.y - p.y, p.y - hi.y), 0.0f);
    float dz = std::fmax(std::fmax(lo.z - p.z, p.z - hi.z), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

inline V3 centroid(const Tri& t) {  // triangle.cuh:220-222
    return v3((t.a.x + t.b.x + t.c.x) / 3.0f, (t.a.y + t.b.y + t.c.y) / 3.0f,
              (t.a.z + t.b.z + t.c.z) / 3.0f);
}

// ---------------------------------------------------------------------------
// Morton codes (tcnn morton3D / morton3D_invert, used at testbed_nerf.cu:325,
// 375-377): bit i of x/y/z goes to bit 3i/3i+1/3i+2.
// ---------------------------------------------------------------------------
inline uint32_t spread3(uint32_t v) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < 10; ++i) r |= ((v >> i) & 1u) << (3 * i);
    return r;
}
inline uint32_t compact3(uint32_t v) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < 10; ++i) r |= ((v >> (3 * i)) & 1u) << i;
    return r;
}
inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

// ---------------------------------------------------------------------------
// BVH.  Declared layout: an implicit complete 4-ary heap (children of node i
// are 4i+1..4i+4) over 4^k count-balanced leaves, k the largest with
// 4^k <= n; at binary depth s the triangles are split into segments
// [floor(j n / 2^s), floor((j+1) n / 2^s)).  Like the reference's build
// (instant-ngp/src/triangle_bvh.cu:548-631) every segment is split at its
// count median along one axis, but the axis is the longest extent of the
// segment's quantised centroids instead of the largest variance, and ties are
// broken by face id, so the layout is an exact function of the input (the
// reference's std::nth_element order is implementation-defined).  Centroids
// are quantised to 18 bits over the centroid AABB; the three per-axis orders
// (q, id) are sorted once and stable-partitioned at every depth.
// Traversal (bvh.cuh:48-96): children sorted far-to-near by box distance with
// a 4-input sorting network, pushed when their distance is <= the best so far
// plus a margin of a few ulp (so the result is the brute-force argmin);
// a leaf triangle replaces the best when (dist_sq, original id) is
// lexicographically smaller -- a layout-independent tie-break instead of the
// reference's "last visited wins" (bvh.cuh:64).
// ---------------------------------------------------------------------------
struct Bvh {
    uint32_t n = 0, k = 0, leaf_base = 0, n_nodes = 0;
    std::vector<V3> lo, hi;
    std::vector<int32_t> begin, end;
    std::vector<Tri> tris;          // BVH order
    std::vector<uint32_t> orig;     // BVH order -> original face id
};

// inner levels of the 4-ary tree: the largest k with 4^k * leaf_min <= n, so
// the 4^k count-balanced leaves hold leaf_min .. 4 * leaf_min - 1 triangles
uint32_t bvh_levels(uint32_t n, uint32_t leaf_min) {
    uint32_t k = 0;
    while (((uint64_t)leaf_min << (2 * (k + 1))) <= n) ++k;
    return k;
}

inline uint32_t quantise18(float c, float lo, float e) {
    if (!(e > 0.0f)) return 0u;
    float q = (c - lo) / e * 262144.0f;
    if (!(q > 0.0f)) return 0u;
    uint32_t u = (uint32_t)q;
    return u > 262143u ? 262143u : u;
}

// segment boundaries at binary depth s
inline uint32_t seg_bound(uint64_t j, uint32_t n, uint32_t s) { return (uint32_t)((j * n) >> s); }

void bvh_order(const std::vector<Tri>& tris, std::vector<uint32_t>& order, uint32_t k) {
    const uint32_t n = (uint32_t)tris.size();
    std::vector<V3> cen(n);
    V3 cmin = v3(INFINITY, INFINITY, INFINITY), cmax = v3(-INFINITY, -INFINITY, -INFINITY);
    for (uint32_t i = 0; i < n; ++i) {
        cen[i] = centroid(tris[i]);
        cmin = vmin(cmin, cen[i]);
        cmax = vmax(cmax, cen[i]);
    }
    const V3 ext = cmax - cmin;
    std::vector<uint32_t> q[3];
    std::vector<uint32_t> list[3];
    for (int a = 0; a < 3; ++a) {
        q[a].resize(n);
        std::vector<uint32_t> keys(n);
        for (uint32_t i = 0; i < n; ++i) {
            const float c = a == 0 ? cen[i].x : (a == 1 ? cen[i].y : cen[i].z);
            const float lo = a == 0 ? cmin.x : (a == 1 ? cmin.y : cmin.z);
            const float e = a == 0 ? ext.x : (a == 1 ? ext.y : ext.z);
            q[a][i] = quantise18(c, lo, e);
            keys[i] = (q[a][i] << 14) | i;
        }
        std::sort(keys.begin(), keys.end());
        list[a].resize(n);
        for (uint32_t i = 0; i < n; ++i) list[a][i] = keys[i] & 0x3fffu;
    }
    std::vector<uint8_t> right(n);
    std::vector<uint32_t> tmp(n);
    for (uint32_t s = 0; s < 2 * k; ++s) {
        const uint32_t n_seg = 1u << s;
        for (uint32_t j = 0; j < n_seg; ++j) {
            const uint32_t b = seg_bound(j, n, s), e = seg_bound(j + 1, n, s);
            const uint32_t mid = seg_bound(2 * j + 1, n, s + 1);
            uint32_t best_axis = 0, best_ext = 0;
            for (uint32_t a = 0; a < 3; ++a) {
                const uint32_t ea = q[a][list[a][e - 1]] - q[a][list[a][b]];
                if (a == 0 || ea > best_ext) {
                    best_ext = ea;
                    best_axis = a;
                }
            }
            for (uint32_t pos = b; pos < e; ++pos) right[list[best_axis][pos]] = pos >= mid ? 1 : 0;
        }
        for (int a = 0; a < 3; ++a) {
            for (uint32_t j = 0; j < n_seg; ++j) {
                const uint32_t b = seg_bound(j, n, s), e = seg_bound(j + 1, n, s);
                uint32_t l = b, r = seg_bound(2 * j + 1, n, s + 1);
                for (uint32_t pos = b; pos < e; ++pos) {
                    const uint32_t id = list[a][pos];
                    tmp[right[id] ? r++ : l++] = id;
                }
            }
            list[a].swap(tmp);
        }
    }
    order = list[0];
}

void bvh_build(Bvh& b, const std::vector<Tri>& tris, uint32_t leaf_min) {
    const uint32_t n = (uint32_t)tris.size();
    b.n = n;
    b.k = bvh_levels(n, leaf_min);
    const uint32_t n_leaves = 1u << (2 * b.k);
    b.leaf_base = (n_leaves - 1) / 3;
    b.n_nodes = b.leaf_base + n_leaves;
    bvh_order(tris, b.orig, b.k);
    b.tris.resize(n);
    for (uint32_t i = 0; i < n; ++i) b.tris[i] = tris[b.orig[i]];
    b.lo.assign(b.n_nodes, v3(0, 0, 0));
    b.hi.assign(b.n_nodes, v3(0, 0, 0));
    b.begin.assign(b.n_nodes, -1);
    b.end.assign(b.n_nodes, -1);
    for (uint32_t j = 0; j < n_leaves; ++j) {
        uint32_t s = (uint32_t)((uint64_t(j) * n) >> (2 * b.k));
        uint32_t e = (uint32_t)((uint64_t(j + 1) * n) >> (2 * b.k));
        V3 lo = v3(INFINITY, INFINITY, INFINITY), hi = v3(-INFINITY, -INFINITY, -INFINITY);
        for (uint32_t t = s; t < e; ++t) {
            lo = vmin(lo, vmin(b.tris[t].a, vmin(b.tris[t].b, b.tris[t].c)));
            hi = vmax(hi, vmax(b.tris[t].a, vmax(b.tris[t].b, b.tris[t].c)));
        }
        uint32_t node = b.leaf_base + j;
        b.lo[node] = lo;
        b.hi[node] = hi;
        b.begin[node] = (int32_t)s;
        b.end[node] = (int32_t)e;
    }
    for (int64_t node = (int64_t)b.leaf_base - 1; node >= 0; --node) {
        V3 lo = v3(INFINITY, INFINITY, INFINITY), hi = v3(-INFINITY, -INFINITY, -INFINITY);
        for (uint32_t c = 0; c < 4; ++c) {
            lo = vmin(lo, b.lo[4 * node + 1 + c]);
            hi = vmax(hi, b.hi[4 * node + 1 + c]);
        }
        b.lo[node] = lo;
        b.hi[node] = hi;
    }
}

struct Hit {
    int32_t idx;      // BVH-order index, -1 if none
    float dist_sq;
};

Hit bvh_closest(const Bvh& b, V3 p) {
    int32_t stack[64];
    int sp = 0;
    stack[sp++] = 0;
    float best = FLT_MAX;
    int32_t best_idx = -1;
    uint32_t best_orig = 0xffffffffu;
    while (sp > 0) {
        int32_t node = stack[--sp];
        if ((uint32_t)node >= b.leaf_base) {
            for (int32_t t = b.begin[node]; t < b.end[node]; ++t) {
                float d = tri_distance_sq(b.tris[t], p);
                uint32_t o = b.orig[t];
                if (d < best || (d == best && o < best_orig)) {
                    best = d;
                    best_idx = t;
                    best_orig = o;
                }
            }
        } else {
            float cd[4];
            int32_t ci[4];
            for (int c = 0; c < 4; ++c) {
                ci[c] = 4 * node + 1 + c;
                cd[c] = box_distance_sq(b.lo[ci[c]], b.hi[ci[c]], p);
            }
            auto cas = [&](int i, int j) {  // larger distance first
                if (cd[i] < cd[j]) {
                    std::swap(cd[i], cd[j]);
                    std::swap(ci[i], ci[j]);
                }
            };
            cas(0, 2); cas(1, 3); cas(0, 1); cas(2, 3); cas(1, 2);
            // conservative prune: a box whose rounded distance exceeds the
            // best rounded triangle distance by a few ulp may still hold the
            // argmin, so keep boxes within a relative 2^-18 + absolute 2^-26
            const float bound = best + best * 0x1p-18f + 0x1p-26f;
            for (int c = 0; c < 4; ++c)
                if (cd[c] <= bound) stack[sp++] = ci[c];
        }
    }
    return {best_idx, best_idx < 0 ? 0.0f : best};
}

// ---------------------------------------------------------------------------
// per-triangle frames, built and inverted in FP64 (north-star precision
// choice); restates Triangle::tbn (triangle.cuh:43-57) and tri2projection
// (include/rta/affine.cuh:26-39).  Stored as row-major 3x4 float [A | t].
// ---------------------------------------------------------------------------
struct Frame12 {
    float m[12];
};

void build_projection(const D3& a, const D3& b, const D3& c, bool scaled, double s, double out[12]) {
    D3 t = b - a;
    D3 bb = c - a;
    D3 n = normalized(crossd(t, bb));
    D3 bi = normalized(crossd(t, n));
    D3 ta = normalized(crossd(bi, n));
    double f = scaled ? s : 1.0;
    D3 cols[3] = {ta, bi, n};
    for (int r = 0; r < 3; ++r) {
        for (int col = 0; col < 3; ++col) {
            double v = r == 0 ? cols[col].x : (r == 1 ? cols[col].y : cols[col].z);
            out[r * 4 + col] = v * f;
        }
    }
    out[3] = a.x;
    out[7] = a.y;
    out[11] = a.z;
}

void invert_affine_d(const double m[12], double out[12]) {
    const double a00 = m[0], a01 = m[1], a02 = m[2];
    const double a10 = m[4], a11 = m[5], a12 = m[6];
    const double a20 = m[8], a21 = m[9], a22 = m[10];
    double c00 = a11 * a22 - a12 * a21, c01 = a02 * a21 - a01 * a22, c02 = a01 * a12 - a02 * a11;
    double c10 = a12 * a20 - a10 * a22, c11 = a00 * a22 - a02 * a20, c12 = a02 * a10 - a00 * a12;
    double c20 = a10 * a21 - a11 * a20, c21 = a01 * a20 - a00 * a21, c22 = a00 * a11 - a01 * a10;
    double det = a00 * c00 + a01 * c10 + a02 * c20;
    double id = 1.0 / det;
    double inv[9] = {c00 * id, c01 * id, c02 * id, c10 * id, c11 * id, c12 * id, c20 * id, c21 * id, c22 * id};
    for (int r = 0; r < 3; ++r) {
        out[r * 4 + 0] = inv[r * 3 + 0];
        out[r * 4 + 1] = inv[r * 3 + 1];
        out[r * 4 + 2] = inv[r * 3 + 2];
        out[r * 4 + 3] = -(inv[r * 3 + 0] * m[3] + inv[r * 3 + 1] * m[7] + inv[r * 3 + 2] * m[11]);
    }
}

// float affine inverse for the per-sample blended frame; the HIP kernel uses
// the same cofactor expansion in the same order.
void invert_affine_f(const float m[12], float out[12]) {
    const float a00 = m[0], a01 = m[1], a02 = m[2];
    const float a10 = m[4], a11 = m[5], a12 = m[6];
    const float a20 = m[8], a21 = m[9], a22 = m[10];
    float c00 = a11 * a22 - a12 * a21, c01 = a02 * a21 - a01 * a22, c02 = a01 * a12 - a02 * a11;
    float c10 = a12 * a20 - a10 * a22, c11 = a00 * a22 - a02 * a20, c12 = a02 * a10 - a00 * a12;
    float c20 = a10 * a21 - a11 * a20, c21 = a01 * a20 - a00 * a21, c22 = a00 * a11 - a01 * a10;
    float det = a00 * c00 + a01 * c10 + a02 * c20;
    float id = 1.0f / det;
    float inv[9] = {c00 * id, c01 * id, c02 * id, c10 * id, c11 * id, c12 * id, c20 * id, c21 * id, c22 * id};
    for (int r = 0; r < 3; ++r) {
        out[r * 4 + 0] = inv[r * 3 + 0];
        out[r * 4 + 1] = inv[r * 3 + 1];
        out[r * 4 + 2] = inv[r * 3 + 2];
        out[r * 4 + 3] = -(inv[r * 3 + 0] * m[3] + inv[r * 3 + 1] * m[7] + inv[r * 3 + 2] * m[11]);
    }
}

inline V3 apply_point(const float m[12], V3 p) {
    return v3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
inline V3 apply_dir(const float m[12], V3 d) {
    return v3(m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
              m[8] * d.x + m[9] * d.y + m[10] * d.z);
}

struct MeshFrame {
    bool built = false;
    std::vector<Tri> tris;        // original order
    std::vector<V3> cen;          // original order, float centroids
    std::vector<Frame12> proj;    // tri2projection
    std::vector<Frame12> proj_inv;
    std::vector<double> proj_d;   // [nf*12] unrounded FP64 frames (FP64 oracle)
    Bvh bvh;
};

// ---------------------------------------------------------------------------
// hash grid layout (declared convention; tiny-cuda-nn GridEncoding is absent
// from the reference tree):  level l has scale s_l = b^l * N_min - 1 (double,
// rounded to float once on the host), resolution r_l = ceil(s_l) + 1, and
// r_l^3 vertices; it is dense (index x + y r + z r^2) when r_l^3 <= T, with
// size rounded up to a multiple of 8, else hashed into T entries with
// h = (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod T (Instant-NGP eq. 4).
// ---------------------------------------------------------------------------
void compute_layout(const insta_config& c, insta_layout& L) {
    std::memset(&L, 0, sizeof(L));
    const uint64_t T = uint64_t(1) << c.log2_hashmap_size;
    uint64_t off = 0;
    for (uint32_t l = 0; l < c.n_levels; ++l) {
        double s = std::exp2(l * std::log2((double)c.per_level_scale)) * c.base_resolution - 1.0;
        float sf = (float)s;
        uint32_t res = (uint32_t)std::ceil(sf) + 1u;
        uint64_t verts = uint64_t(res) * res * res;
        bool dense = verts <= T;
        uint64_t size = dense ? ((verts + 7) / 8) * 8 : T;
        L.level_scale[l] = sf;
        L.level_resolution[l] = res;
        L.level_dense[l] = dense ? 1u : 0u;
        L.level_offset[l] = (uint32_t)off;
        L.level_size[l] = (uint32_t)size;
        off += size;
    }
    L.n_grid_params = off * c.n_features_per_level;
    L.density_in = c.n_levels * c.n_features_per_level + c.n_extra_dims;
    L.rgb_in = c.density_out + c.sh_degree * c.sh_degree;
    uint64_t w = c.n_neurons;
    uint64_t n_mlp = w * L.density_in + (uint64_t)(c.density_hidden_layers - 1) * w * w + c.density_out * w;
    n_mlp += w * L.rgb_in + (uint64_t)(c.rgb_hidden_layers - 1) * w * w + 3 * w;
    L.n_mlp_params = n_mlp;
    L.n_params = n_mlp + L.n_grid_params;
}

inline uint32_t grid_index(const insta_layout& L, uint32_t l, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t size = L.level_size[l];
    if (L.level_dense[l]) {
        const uint32_t r = L.level_resolution[l];
        return (x + y * r + z * (r * r)) % size;
    }
    return (x ^ (y * 2654435761u) ^ (z * 805459861u)) % size;
}

// ---------------------------------------------------------------------------
// Sobol + Owen scrambling (random_val.cuh:159-288), after Burley 2020,
// "Practical Hash-based Owen Scrambling" (JCGT 9(4)).  Dimension 0 is the
// van der Corput sequence (bit reversal); dimension 1 uses direction numbers
// v_0 = 2^31, v_k = v_{k-1} ^ (v_{k-1} >> 1).
// ---------------------------------------------------------------------------
inline uint32_t reverse_bits(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}
inline uint32_t sobol_dim(uint32_t index, uint32_t dim) {
    if (dim == 0) return reverse_bits(index);
    uint32_t v = 0x80000000u, x = 0;
    for (uint32_t bit = 0; bit < 32; ++bit) {
        if ((index >> bit) & 1u) x ^= v;
        v ^= v >> 1;
    }
    return x;
}
inline uint32_t hash_combine(uint32_t seed, uint32_t v) { return seed ^ (v + (seed << 6) + (seed >> 2)); }
inline uint32_t lk_permutation(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
inline uint32_t owen_scramble(uint32_t x, uint32_t seed) {
    return reverse_bits(lk_permutation(reverse_bits(x), seed));
}
inline float ld_random_val(uint32_t index, uint32_t seed, uint32_t dim) {
    const float S = (float)(1.0 / 4294967296.0);
    index = owen_scramble(index, seed);
    return (float)owen_scramble(sobol_dim(index, dim), hash_combine(seed, dim)) * S;
}
inline void ld_random_val_2d(uint32_t index, uint32_t seed, float out[2]) {
    const float S = (float)(1.0 / 4294967296.0);
    index = owen_scramble(index, seed);
    for (uint32_t i = 0; i < 2; ++i) out[i] = (float)owen_scramble(sobol_dim(index, i), hash_combine(seed, i)) * S;
}
// random_val.cuh:317-322
inline void ld_random_pixel_offset(uint32_t spp, float out[2]) {
    float a[2], b[2];
    ld_random_val_2d(0, 0xdeadbeefu, a);
    ld_random_val_2d(spp, 0xdeadbeefu, b);
    for (int i = 0; i < 2; ++i) {
        float o = 0.5f - a[i] + b[i];
        out[i] = o - std::floor(o);
    }
}

// ---------------------------------------------------------------------------
// real spherical harmonics, degree 4 = bands 0..3 (16 coefficients) of the
// direction mapped from [0,1]^3 to [-1,1]^3 (tcnn SphericalHarmonics encoding,
// configs/nerf/main.json:48-62).  Constants are the orthonormal real SH
// normalisations: 1/(2 sqrt(pi)), sqrt(3/(4 pi)), sqrt(15/pi)/2, sqrt(5/pi)/4,
// sqrt(15/pi)/4, sqrt(35/(2 pi))/4, sqrt(105/pi)/2, sqrt(21/(2 pi))/4,
// sqrt(7/pi)/4, sqrt(105/pi)/4.
// ---------------------------------------------------------------------------
void sh16(float x, float y, float z, float* o) {
    const float xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
    o[0] = 0.28209479177387814f;
    o[1] = (ORC_MUTANT == 2 ? 0.48860251190291987f : -0.48860251190291987f) * y;
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

inline float logistic(float x) { return 1.0f / (1.0f + std::exp(-x)); }

}  // namespace

// ===========================================================================
// scene (mirror of insta_ctx)
// ===========================================================================
struct orc_scene {
    insta_config cfg{};
    insta_layout layout{};
    uint32_t nv = 0, nf = 0;
    std::vector<int32_t> faces;
    std::vector<uint32_t> masks;
    std::vector<int32_t> adjacency;   // [nf*3]
    MeshFrame canon;
    std::vector<MeshFrame> frames;
    V3 render_lo{}, render_hi{};
    std::vector<float> params;
    std::vector<std::vector<uint8_t>> grids;   // slot 0 = canonical, slot f+1 = frame f
    int n_threads = 0;
    // density-grid EMA (orc_update_density_grid)
    std::vector<float> dg_grid;
    uint32_t dg_step = 0;
    uint32_t dg_marked_n = 0;     // training images at the last untrained-cell marking
    uint64_t dg_rng_state = 0, dg_rng_inc = 0;
    // training set (orc_set_training_set)
    uint32_t tr_n = 0, tr_w = 0, tr_h = 0;
    std::vector<float> tr_images;              // [n][h][w][4]
    std::vector<insta_camera> tr_cams;
    std::vector<int32_t> tr_frames;
    int32_t tr_occ_slot = -2;     // >= -1: every training ray marches this slot's grid
    std::vector<uint8_t> tr_seg;  // [n][h][w] mask ids, empty: no masks (orc_set_training_supervision)
    std::vector<float> tr_depth;  // [n][h][w] camera-axis depth, 0 = none; empty: no depth images
    double tr_seg_weight = 1.0, tr_depth_weight = 1.5;
    std::vector<float> frame_extra;  // [max_frames][n_extra_dims] (Core::load_meshes' extra_dims_gpu)
};

namespace {

void fill_frame(orc_scene* s, MeshFrame& mf, const float* verts, const MeshFrame* canon) {
    const uint32_t nf = s->nf;
    mf.tris.resize(nf);
    mf.cen.resize(nf);
    mf.proj.resize(nf);
    mf.proj_inv.resize(nf);
    mf.proj_d.resize(12 * (size_t)nf);
    for (uint32_t i = 0; i < nf; ++i) {
        const int32_t* f = &s->faces[3 * i];
        Tri t;
        t.a = v3(verts[3 * f[0]], verts[3 * f[0] + 1], verts[3 * f[0] + 2]);
        t.b = v3(verts[3 * f[1]], verts[3 * f[1] + 1], verts[3 * f[1] + 2]);
        t.c = v3(verts[3 * f[2]], verts[3 * f[2] + 1], verts[3 * f[2] + 2]);
        mf.tris[i] = t;
        mf.cen[i] = centroid(t);
        D3 a = {t.a.x, t.a.y, t.a.z}, b = {t.b.x, t.b.y, t.b.z}, c = {t.c.x, t.c.y, t.c.z};
        double s_ratio = 1.0;
        if (canon) {
            const Tri& ct = canon->tris[i];
            D3 ca = {ct.a.x, ct.a.y, ct.a.z}, cb = {ct.b.x, ct.b.y, ct.b.z}, cc = {ct.c.x, ct.c.y, ct.c.z};
            D3 cr = crossd(b - a, c - a), ccr = crossd(cb - ca, cc - ca);
            double area = 0.5 * std::sqrt(dotd(cr, cr)), carea = 0.5 * std::sqrt(dotd(ccr, ccr));
            s_ratio = area / carea;
        }
        double P[12], Pi[12];
        build_projection(a, b, c, s->masks[i] == 1u, s_ratio, P);
        invert_affine_d(P, Pi);
        for (int k = 0; k < 12; ++k) {
            mf.proj[i].m[k] = (float)P[k];
            mf.proj_inv[i].m[k] = (float)Pi[k];
            mf.proj_d[12 * (size_t)i + k] = P[k];
        }
    }
    bvh_build(mf.bvh, mf.tris, s->cfg.bvh_leaf_min);
    mf.built = true;
}

// interpolate_transformation (include/rta/affine.cuh:45-100) without the
// final inverse: exp-weighted blend of the triangle's own frame (factor 1)
// and its edge neighbours' frames (factor 4), weights normalised.
constexpr float kBlendSelf = ORC_MUTANT == 1 ? 4.0f : 1.0f;       // affine.cuh:48
constexpr float kBlendNeighbour = ORC_MUTANT == 1 ? 1.0f : 4.0f;  // affine.cuh:56, 60
void blend_frames(const orc_scene* s, const MeshFrame& mf, uint32_t id, V3 p, float out[12]) {
    const int32_t* adj = &s->adjacency[3 * id];
    float sum = 0.0f;
    for (int k = 0; k < 3; ++k)
        if (adj[k] >= 0) sum += std::exp(-kBlendNeighbour * norm(mf.cen[adj[k]] - p));
    sum += std::exp(-kBlendSelf * norm(mf.cen[id] - p));
    for (int i = 0; i < 12; ++i) out[i] = 0.0f;
    for (int k = 0; k < 3; ++k) {
        if (adj[k] >= 0) {
            float a = std::exp(-kBlendNeighbour * norm(mf.cen[adj[k]] - p));
            float w = a / sum;
            const float* m = mf.proj[adj[k]].m;
            for (int i = 0; i < 12; ++i) out[i] = out[i] + w * m[i];
        }
    }
    float a = std::exp(-kBlendSelf * norm(mf.cen[id] - p));
    float w = a / sum;
    const float* m = mf.proj[id].m;
    for (int i = 0; i < 12; ++i) out[i] = out[i] + w * m[i];
}

// The per-sample body of `warp` (src/core.cu:470-519).  Returns false when the
// sample is left untouched.
bool warp_one(const orc_scene* s, const MeshFrame& fr, V3 p, V3 dir_w, V3& out_p, V3& out_d, uint32_t& mask) {
    if (!(p.x >= 0.0f && p.x <= 1.0f && p.y >= 0.0f && p.y <= 1.0f && p.z >= 0.0f && p.z <= 1.0f)) return false;
    Hit h = bvh_closest(fr.bvh, p);
    if (h.idx < 0) return false;
    const uint32_t id = fr.bvh.orig[h.idx];
    mask = s->masks[id];
    float W[12], C[12];
    if (mask == 2u && !s->cfg.disable_interp) {
        float Wb[12];
        blend_frames(s, fr, id, p, Wb);
        invert_affine_f(Wb, W);
        blend_frames(s, s->canon, id, p, C);
    } else {
        std::memcpy(W, fr.proj_inv[id].m, sizeof(W));
        std::memcpy(C, s->canon.proj[id].m, sizeof(C));
    }
    out_p = apply_point(C, apply_point(W, p));
    V3 d = v3(dir_w.x * 2.0f - 1.0f, dir_w.y * 2.0f - 1.0f, dir_w.z * 2.0f - 1.0f);
    V3 dc = apply_dir(C, apply_dir(W, d));
    float n = norm(dc);
    dc = v3(dc.x / n, dc.y / n, dc.z / n);
    out_d = v3((dc.x + 1.0f) * 0.5f, (dc.y + 1.0f) * 0.5f, (dc.z + 1.0f) * 0.5f);
    return true;
}

const MeshFrame* frame_ptr(const orc_scene* s, int32_t frame) {
    if (frame < 0) return s->canon.built ? &s->canon : nullptr;
    if ((uint32_t)frame >= s->frames.size() || !s->frames[frame].built) return nullptr;
    return &s->frames[frame];
}

// hash-grid forward for one position (see the declared convention above)
void encode_one(const orc_scene* s, V3 x, float* out) {
    const insta_layout& L = s->layout;
    const uint32_t F = s->cfg.n_features_per_level;
    const float* table = s->params.data() + L.n_mlp_params;
    const float xs[3] = {x.x, x.y, x.z};
    for (uint32_t l = 0; l < s->cfg.n_levels; ++l) {
        const float sc = L.level_scale[l];
        float frac[3];
        uint32_t g[3];
        for (int d = 0; d < 3; ++d) {
            float pos = xs[d] * sc + 0.5f;
            float fl = std::floor(pos);
            frac[d] = pos - fl;
            g[d] = (uint32_t)(int32_t)fl;
        }
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t c = 0; c < 8; ++c) {
            float w = 1.0f;
            uint32_t cg[3];
            for (int d = 0; d < 3; ++d) {
                bool up = (c >> d) & 1u;
                w = w * (up ? frac[d] : 1.0f - frac[d]);
                cg[d] = g[d] + (up ? 1u : 0u);
            }
            uint32_t idx = grid_index(L, l, cg[0], cg[1], cg[2]);
            const float* e = table + (uint64_t(L.level_offset[l]) + idx) * F;
            for (uint32_t f = 0; f < F; ++f) acc[f] = acc[f] + w * e[f];
        }
        for (uint32_t f = 0; f < F; ++f) out[l * F + f] = acc[f];
    }
}

// bias-free dense layer, row-major W[out][in], sequential accumulation
void dense(const float* W, const float* x, uint32_t n_in, uint32_t n_out, bool relu, float* y) {
    for (uint32_t o = 0; o < n_out; ++o) {
        float acc = 0.0f;
        const float* w = W + (uint64_t)o * n_in;
        for (uint32_t i = 0; i < n_in; ++i) acc = acc + w[i] * x[i];
        y[o] = relu ? (acc > 0.0f ? acc : 0.0f) : acc;
    }
}

// NerfNetwork::inference_mixed_precision_impl (nerf_network.h:159-232) for
// one sample: encoding [+ extra dims] -> density MLP -> [density out | SH]
// -> rgb MLP.  out4 = rgb logits, density logit (density output 0).
void network_one(const orc_scene* s, V3 pos, V3 dir_w, const float* extra, float* out4) {
    const insta_config& c = s->cfg;
    const insta_layout& L = s->layout;
    const uint32_t W = c.n_neurons;
    float din[INSTA_MAX_LEVELS * 8 + 64];
    float h0[256], h1[256], rgb_in[64];
    encode_one(s, pos, din);
    const uint32_t enc = c.n_levels * c.n_features_per_level;
    for (uint32_t e = 0; e < c.n_extra_dims; ++e) din[enc + e] = extra ? extra[e] : 0.0f;
    const float* p = s->params.data();
    dense(p, din, L.density_in, W, true, h0);
    p += (uint64_t)W * L.density_in;
    for (uint32_t k = 1; k < c.density_hidden_layers; ++k) {
        dense(p, h0, W, W, true, h1);
        std::memcpy(h0, h1, W * sizeof(float));
        p += (uint64_t)W * W;
    }
    dense(p, h0, W, c.density_out, false, rgb_in);
    p += (uint64_t)c.density_out * W;
    float dx = dir_w.x * 2.0f - 1.0f, dy = dir_w.y * 2.0f - 1.0f, dz = dir_w.z * 2.0f - 1.0f;
    sh16(dx, dy, dz, rgb_in + c.density_out);
    dense(p, rgb_in, L.rgb_in, W, true, h0);
    p += (uint64_t)W * L.rgb_in;
    for (uint32_t k = 1; k < c.rgb_hidden_layers; ++k) {
        dense(p, h0, W, W, true, h1);
        std::memcpy(h0, h1, W * sizeof(float));
        p += (uint64_t)W * W;
    }
    float rgb[3];
    dense(p, h0, W, 3, false, rgb);
    out4[0] = rgb[0];
    out4[1] = rgb[1];
    out4[2] = rgb[2];
    out4[3] = rgb_in[0];
}

// Extra dims of a sample: the frame's values (set per sample by
// generate_next_nerf_network_inputs / generate_training_samples_nerf,
// testbed_nerf.cu:788, 1158, 1270) unless warp deformed it off the lips
// region, which writes 1.0 (src/core.cu:514-519); no frame: 1.0.
const float* sample_extra(const orc_scene* s, int32_t frame, bool warped, uint32_t mask) {
    static const std::vector<float> ones(64, 1.0f);
    if (frame < 0 || (warped && mask != 7u)) return ones.data();
    return s->frame_extra.data() + (size_t)frame * s->cfg.n_extra_dims;
}

// ---------------------------------------------------------------------------
// occupancy + march helpers (testbed_nerf.cu:186-208, 313-337, 634-686, 722-795)
// ---------------------------------------------------------------------------
inline bool occupied(const uint8_t* grid, uint32_t G, V3 pos) {
    // cascaded_grid_idx_at with mip 0: (pos - 0.5) * 1 + 0.5, scaled by G, truncated
    float px = ((pos.x - 0.5f) * 1.0f + 0.5f) * (float)G;
    float py = ((pos.y - 0.5f) * 1.0f + 0.5f) * (float)G;
    float pz = ((pos.z - 0.5f) * 1.0f + 0.5f) * (float)G;
    int ix = (int)px, iy = (int)py, iz = (int)pz;
    const int hi = (int)G - 1;
    ix = ix < 0 ? 0 : (ix > hi ? hi : ix);
    iy = iy < 0 ? 0 : (iy > hi ? hi : iy);
    iz = iz < 0 ? 0 : (iz > hi ? hi : iz);
    uint32_t idx = morton3((uint32_t)ix, (uint32_t)iy, (uint32_t)iz);
    return (grid[idx / 8] >> (idx % 8)) & 1u;
}

inline bool box_contains(V3 lo, V3 hi, V3 p) {
    return p.x >= lo.x && p.x <= hi.x && p.y >= lo.y && p.y <= hi.y && p.z >= lo.z && p.z <= hi.z;
}

inline float distance_to_next_voxel(V3 pos, V3 dir, V3 idir, uint32_t res) {
    const float r = (float)res;
    V3 p = v3(r * pos.x, r * pos.y, r * pos.z);
    float tx = (std::floor(p.x + 0.5f + 0.5f * signf(dir.x)) - p.x) * idir.x;
    float ty = (std::floor(p.y + 0.5f + 0.5f * signf(dir.y)) - p.y) * idir.y;
    float tz = (std::floor(p.z + 0.5f + 0.5f * signf(dir.z)) - p.z) * idir.z;
    float t = std::fmin(std::fmin(tx, ty), tz);
    return std::fmax(t / r, 0.0f);
}

inline float advance_to_next_voxel(float t, float dt, V3 pos, V3 dir, V3 idir, uint32_t res) {
    float t_target = t + distance_to_next_voxel(pos, dir, idir, res);
    do {
        t += dt;
    } while (t < t_target);
    return t;
}

// BoundingBox::ray_intersect (bounding_box.cuh:163-210), returns tmin
inline float box_ray_tmin(V3 lo, V3 hi, V3 o, V3 d) {
    float tmin = (lo.x - o.x) / d.x, tmax = (hi.x - o.x) / d.x;
    if (tmin > tmax) std::swap(tmin, tmax);
    float tymin = (lo.y - o.y) / d.y, tymax = (hi.y - o.y) / d.y;
    if (tymin > tymax) std::swap(tymin, tymax);
    if (tmin > tymax || tymin > tmax) return FLT_MAX;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (lo.z - o.z) / d.z, tzmax = (hi.z - o.z) / d.z;
    if (tzmin > tzmax) std::swap(tzmin, tzmax);
    if (tmin > tzmax || tzmin > tmax) return FLT_MAX;
    if (tzmin > tmin) tmin = tzmin;
    return tmin;
}

// ---------------------------------------------------------------------------
// FP64 restatement of the per-sample math (deform, encode, MLPs, composite),
// used as the accuracy reference for the FP32/FP16/BF16 renders.  The march
// itself (sample positions) stays FP32 so all paths see the same samples.
// ---------------------------------------------------------------------------
struct D3v {
    double x, y, z;
};
inline D3v dv(V3 a) { return {a.x, a.y, a.z}; }
inline D3v operator-(D3v a, D3v b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3v operator*(D3v a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double ddot(D3v a, D3v b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3v dcross(D3v a, D3v b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

double tri_distance_sq_d(const Tri& t, D3v pos) {
    D3v a = dv(t.a), b = dv(t.b), c = dv(t.c);
    D3v v21 = b - a, p1 = pos - a, v32 = c - b, p2 = pos - b, v13 = a - c, p3 = pos - c;
    D3v nor = dcross(v21, v13);
    double sg = std::copysign(1.0, ddot(dcross(v21, nor), p1)) + std::copysign(1.0, ddot(dcross(v32, nor), p2)) +
                std::copysign(1.0, ddot(dcross(v13, nor), p3));
    if (sg < 2.0) {
        auto edge = [](D3v v, D3v p) {
            double t = std::fmin(std::fmax(ddot(v, p) / ddot(v, v), 0.0), 1.0);
            D3v d = v * t - p;
            return ddot(d, d);
        };
        return std::fmin(std::fmin(edge(v21, p1), edge(v32, p2)), edge(v13, p3));
    }
    double d = ddot(nor, p1);
    return d * d / ddot(nor, nor);
}

// nearest triangle in FP64 over the same BVH (boxes are exact float bounds)
int32_t bvh_closest_d(const Bvh& b, D3v p) {
    int32_t stack[64];
    int sp = 0;
    stack[sp++] = 0;
    double best = 1e300;
    int32_t bi = -1;
    uint32_t bo = 0xffffffffu;
    while (sp > 0) {
        int32_t node = stack[--sp];
        if ((uint32_t)node >= b.leaf_base) {
            for (int32_t t = b.begin[node]; t < b.end[node]; ++t) {
                double d = tri_distance_sq_d(b.tris[t], p);
                if (d < best || (d == best && b.orig[t] < bo)) {
                    best = d;
                    bi = t;
                    bo = b.orig[t];
                }
            }
        } else {
            for (int c = 0; c < 4; ++c) {
                const int32_t ch = 4 * node + 1 + c;
                double dx = std::fmax(std::fmax(b.lo[ch].x - p.x, p.x - b.hi[ch].x), 0.0);
                double dy = std::fmax(std::fmax(b.lo[ch].y - p.y, p.y - b.hi[ch].y), 0.0);
                double dz = std::fmax(std::fmax(b.lo[ch].z - p.z, p.z - b.hi[ch].z), 0.0);
                if (dx * dx + dy * dy + dz * dz <= best * (1.0 + 1e-12)) stack[sp++] = ch;
            }
        }
    }
    return bi;
}

void blend_frames_d(const orc_scene* s, const MeshFrame& mf, uint32_t id, D3v p, double out[12]) {
    const int32_t* adj = &s->adjacency[3 * id];
    auto cen = [&](uint32_t t) { return dv(mf.cen[t]); };
    auto nrm = [](D3v a) { return std::sqrt(ddot(a, a)); };
    double sum = 0.0;
    for (int k = 0; k < 3; ++k)
        if (adj[k] >= 0) sum += std::exp(-4.0 * nrm(cen(adj[k]) - p));
    sum += std::exp(-1.0 * nrm(cen(id) - p));
    for (int i = 0; i < 12; ++i) out[i] = 0.0;
    for (int k = 0; k < 3; ++k) {
        if (adj[k] >= 0) {
            double w = std::exp(-4.0 * nrm(cen(adj[k]) - p)) / sum;
            for (int i = 0; i < 12; ++i) out[i] += w * mf.proj_d[12 * (size_t)adj[k] + i];
        }
    }
    double w = std::exp(-1.0 * nrm(cen(id) - p)) / sum;
    for (int i = 0; i < 12; ++i) out[i] += w * mf.proj_d[12 * (size_t)id + i];
}

inline D3v apply_point_d(const double m[12], D3v p) {
    return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
inline D3v apply_dir_d(const double m[12], D3v d) {
    return {m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
            m[8] * d.x + m[9] * d.y + m[10] * d.z};
}

bool warp_one_d(const orc_scene* s, const MeshFrame& fr, D3v p, D3v dir, D3v& op, D3v& od, uint32_t& mask) {
    if (!(p.x >= 0.0 && p.x <= 1.0 && p.y >= 0.0 && p.y <= 1.0 && p.z >= 0.0 && p.z <= 1.0)) return false;
    int32_t idx = bvh_closest_d(fr.bvh, p);
    if (idx < 0) return false;
    const uint32_t id = fr.bvh.orig[idx];
    mask = s->masks[id];
    double W[12], C[12];
    if (mask == 2u && !s->cfg.disable_interp) {
        double Wb[12];
        blend_frames_d(s, fr, id, p, Wb);
        invert_affine_d(Wb, W);
        blend_frames_d(s, s->canon, id, p, C);
    } else {
        invert_affine_d(&fr.proj_d[12 * (size_t)id], W);
        std::memcpy(C, &s->canon.proj_d[12 * (size_t)id], sizeof(C));
    }
    op = apply_point_d(C, apply_point_d(W, p));
    D3v dc = apply_dir_d(C, apply_dir_d(W, dir));
    od = dc * (1.0 / std::sqrt(ddot(dc, dc)));
    return true;
}

void encode_one_d(const orc_scene* s, D3v x, double* out) {
    const insta_layout& L = s->layout;
    const uint32_t F = s->cfg.n_features_per_level;
    const float* table = s->params.data() + L.n_mlp_params;
    const double xs[3] = {x.x, x.y, x.z};
    for (uint32_t l = 0; l < s->cfg.n_levels; ++l) {
        double frac[3];
        uint32_t g[3];
        for (int d = 0; d < 3; ++d) {
            double pos = xs[d] * (double)L.level_scale[l] + 0.5;
            double fl = std::floor(pos);
            frac[d] = pos - fl;
            g[d] = (uint32_t)(int32_t)fl;
        }
        for (uint32_t f = 0; f < F; ++f) out[l * F + f] = 0.0;
        for (uint32_t c = 0; c < 8; ++c) {
            double w = 1.0;
            for (int d = 0; d < 3; ++d) w *= ((c >> d) & 1u) ? frac[d] : 1.0 - frac[d];
            uint32_t idx = grid_index(L, l, g[0] + (c & 1u), g[1] + ((c >> 1) & 1u), g[2] + ((c >> 2) & 1u));
            const float* e = table + (uint64_t(L.level_offset[l]) + idx) * F;
            for (uint32_t f = 0; f < F; ++f) out[l * F + f] += w * (double)e[f];
        }
    }
}

void dense_d(const float* W, const double* x, uint32_t n_in, uint32_t n_out, bool relu, double* y) {
    for (uint32_t o = 0; o < n_out; ++o) {
        double acc = 0.0;
        for (uint32_t i = 0; i < n_in; ++i) acc += (double)W[(uint64_t)o * n_in + i] * x[i];
        y[o] = relu ? (acc > 0.0 ? acc : 0.0) : acc;
    }
}

void network_one_d(const orc_scene* s, D3v pos, D3v dir, const float* extra, double* out4) {
    const insta_config& c = s->cfg;
    const insta_layout& L = s->layout;
    const uint32_t W = c.n_neurons;
    double din[INSTA_MAX_LEVELS * 8 + 64] = {0};
    double h0[256], h1[256], rgb_in[64];
    encode_one_d(s, pos, din);
    const uint32_t enc = c.n_levels * c.n_features_per_level;
    for (uint32_t e = 0; e < c.n_extra_dims; ++e) din[enc + e] = extra[e];
    const float* p = s->params.data();
    dense_d(p, din, L.density_in, W, true, h0);
    p += (uint64_t)W * L.density_in;
    for (uint32_t k = 1; k < c.density_hidden_layers; ++k) {
        dense_d(p, h0, W, W, true, h1);
        std::memcpy(h0, h1, W * sizeof(double));
        p += (uint64_t)W * W;
    }
    dense_d(p, h0, W, c.density_out, false, rgb_in);
    p += (uint64_t)c.density_out * W;
    const double x = dir.x, y = dir.y, z = dir.z;
    double* o = rgb_in + c.density_out;
    const double pi = 3.14159265358979323846;
    const double c0 = 0.5 / std::sqrt(pi), c1 = std::sqrt(3.0 / (4 * pi)), c2 = 0.5 * std::sqrt(15.0 / pi),
                 c3 = 0.25 * std::sqrt(5.0 / pi), c4 = 0.25 * std::sqrt(15.0 / pi), c5 = 0.25 * std::sqrt(35.0 / (2 * pi)),
                 c6 = 0.5 * std::sqrt(105.0 / pi), c7 = 0.25 * std::sqrt(21.0 / (2 * pi)), c8 = 0.25 * std::sqrt(7.0 / pi),
                 c9 = 0.25 * std::sqrt(105.0 / pi);
    o[0] = c0; o[1] = -c1 * y; o[2] = c1 * z; o[3] = -c1 * x;
    o[4] = c2 * x * y; o[5] = -c2 * y * z; o[6] = c3 * (3 * z * z - 1); o[7] = -c2 * x * z;
    o[8] = c4 * (x * x - y * y); o[9] = c5 * y * (3 * x * x - y * y) * -1.0; o[10] = c6 * x * y * z;
    o[11] = c7 * y * (1 - 5 * z * z); o[12] = c8 * z * (5 * z * z - 3); o[13] = c7 * x * (1 - 5 * z * z);
    o[14] = c9 * z * (x * x - y * y); o[15] = c5 * x * (-x * x + 3 * y * y);
    dense_d(p, rgb_in, L.rgb_in, W, true, h0);
    p += (uint64_t)W * L.rgb_in;
    for (uint32_t k = 1; k < c.rgb_hidden_layers; ++k) {
        dense_d(p, h0, W, W, true, h1);
        std::memcpy(h0, h1, W * sizeof(double));
        p += (uint64_t)W * W;
    }
    double rgb[3];
    dense_d(p, h0, W, 3, false, rgb);
    out4[0] = rgb[0];
    out4[1] = rgb[1];
    out4[2] = rgb[2];
    out4[3] = rgb_in[0];
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

uint32_t orc_abi_version(void) { return 1u; }

orc_scene* orc_create(const insta_config* cfg) {
    if (!cfg || cfg->n_levels == 0 || cfg->n_levels > INSTA_MAX_LEVELS) return nullptr;
    orc_scene* s = new orc_scene();
    s->cfg = *cfg;
    compute_layout(s->cfg, s->layout);
    s->frames.resize(cfg->max_frames);
    s->grids.resize(cfg->max_frames + 1);
    s->frame_extra.assign((size_t)cfg->max_frames * cfg->n_extra_dims, 0.0f);
    return s;
}

void orc_destroy(orc_scene* s) { delete s; }

void orc_set_threads(orc_scene* s, int n) { s->n_threads = n; }

int orc_query_layout(const insta_config* cfg, insta_layout* out) {
    compute_layout(*cfg, *out);
    return 0;
}

int orc_set_canonical_mesh(orc_scene* s, const float* verts, uint32_t nv, const int32_t* faces, uint32_t nf,
                           const uint32_t* masks) {
    s->nv = nv;
    s->nf = nf;
    s->faces.assign(faces, faces + 3 * (size_t)nf);
    s->masks.assign(masks, masks + nf);
    // get_triangle_3_neighbours (src/tiny_mesh.cu:316-371): faces sharing an
    // edge, ascending face id, self removed, -1 padded
    std::vector<std::vector<int32_t>> vf(nv);
    for (uint32_t i = 0; i < nf; ++i)
        for (int k = 0; k < 3; ++k) vf[faces[3 * i + k]].push_back((int32_t)i);
    s->adjacency.assign(3 * (size_t)nf, -1);
    for (uint32_t i = 0; i < nf; ++i) {
        std::vector<int32_t> ids;
        const int32_t v[3] = {faces[3 * i], faces[3 * i + 1], faces[3 * i + 2]};
        const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (auto& pr : pairs) {
            const auto& A = vf[v[pr[0]]];
            const auto& B = vf[v[pr[1]]];
            for (int32_t f : A)
                if (std::find(B.begin(), B.end(), f) != B.end()) ids.push_back(f);
        }
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        int k = 0;
        for (int32_t f : ids) {
            if (f == (int32_t)i) continue;
            if (k < 3) s->adjacency[3 * i + k] = f;
            ++k;
        }
    }
    fill_frame(s, s->canon, verts, nullptr);
    // render box: vertex AABB inflated by 0.15 |diag| (tiny_mesh.cu:157-162),
    // intersected with the unit cube (testbed_nerf.cu:2675-2682)
    V3 lo = v3(INFINITY, INFINITY, INFINITY), hi = v3(-INFINITY, -INFINITY, -INFINITY);
    for (uint32_t i = 0; i < nv; ++i) {
        V3 p = v3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    float amount = norm(hi - lo) * 0.15f;
    lo = lo - v3(amount, amount, amount);
    hi = hi + v3(amount, amount, amount);
    s->render_lo = vmax(lo, v3(0.0f, 0.0f, 0.0f));
    s->render_hi = vmin(hi, v3(1.0f, 1.0f, 1.0f));
    return 0;
}

int orc_set_frame_vertices(orc_scene* s, uint32_t frame, const float* verts) {
    if (frame >= s->frames.size() || !s->canon.built) return 1;
    fill_frame(s, s->frames[frame], verts, &s->canon);
    return 0;
}

int orc_set_frame_extra_dims(orc_scene* s, uint32_t first, uint32_t n, const float* values) {
    const uint32_t X = s->cfg.n_extra_dims;
    if (X == 0 || (uint64_t)first + n > s->frames.size()) return 1;
    std::memcpy(s->frame_extra.data() + (size_t)first * X, values, (size_t)n * X * sizeof(float));
    return 0;
}

void orc_render_box(const orc_scene* s, float* lo_hi) {
    lo_hi[0] = s->render_lo.x; lo_hi[1] = s->render_lo.y; lo_hi[2] = s->render_lo.z;
    lo_hi[3] = s->render_hi.x; lo_hi[4] = s->render_hi.y; lo_hi[5] = s->render_hi.z;
}

void orc_adjacency(const orc_scene* s, int32_t* out) {
    std::memcpy(out, s->adjacency.data(), s->adjacency.size() * sizeof(int32_t));
}

// per-triangle frames of a frame (frame = -1: canonical), [nf*12] each
int orc_frames(const orc_scene* s, int32_t frame, float* proj, float* proj_inv, float* centroids) {
    const MeshFrame* fr = frame_ptr(s, frame);
    if (!fr) return 1;
    for (uint32_t i = 0; i < s->nf; ++i) {
        if (proj) std::memcpy(proj + 12 * i, fr->proj[i].m, 48);
        if (proj_inv) std::memcpy(proj_inv + 12 * i, fr->proj_inv[i].m, 48);
        if (centroids) {
            centroids[3 * i] = fr->cen[i].x;
            centroids[3 * i + 1] = fr->cen[i].y;
            centroids[3 * i + 2] = fr->cen[i].z;
        }
    }
    return 0;
}

uint32_t orc_bvh_n_nodes(const orc_scene* s, int32_t frame) {
    const MeshFrame* fr = frame_ptr(s, frame);
    return fr ? fr->bvh.n_nodes : 0u;
}

int orc_export_bvh(const orc_scene* s, int32_t frame, float* nodes, uint32_t* order) {
    const MeshFrame* fr = frame_ptr(s, frame);
    if (!fr) return 1;
    const Bvh& b = fr->bvh;
    for (uint32_t i = 0; i < b.n_nodes; ++i) {
        float* o = nodes + 8 * i;
        o[0] = b.lo[i].x; o[1] = b.lo[i].y; o[2] = b.lo[i].z;
        std::memcpy(&o[3], &b.begin[i], 4);
        o[4] = b.hi[i].x; o[5] = b.hi[i].y; o[6] = b.hi[i].z;
        std::memcpy(&o[7], &b.end[i], 4);
    }
    if (order) std::memcpy(order, b.orig.data(), b.n * sizeof(uint32_t));
    return 0;
}

int orc_closest_triangles(const orc_scene* s, int32_t frame, const float* pts, uint32_t n, int32_t* ids, float* dist) {
    const MeshFrame* fr = frame_ptr(s, frame);
    if (!fr) return 1;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        Hit h = bvh_closest(fr->bvh, v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
        ids[i] = h.idx < 0 ? -1 : (int32_t)fr->bvh.orig[h.idx];
        if (dist) dist[i] = std::sqrt(h.dist_sq);
    }
    return 0;
}

// brute force over all triangles with the same (dist_sq, id) order: the
// pinning check of the BVH traversal
int orc_closest_triangles_brute(const orc_scene* s, int32_t frame, const float* pts, uint32_t n, int32_t* ids,
                                float* dist_sq) {
    const MeshFrame* fr = frame_ptr(s, frame);
    if (!fr) return 1;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        V3 p = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        float best = FLT_MAX;
        int32_t bi = -1;
        for (uint32_t t = 0; t < s->nf; ++t) {
            float d = tri_distance_sq(fr->tris[t], p);
            if (d < best) {
                best = d;
                bi = (int32_t)t;
            }
        }
        ids[i] = bi;
        dist_sq[i] = best;
    }
    return 0;
}

float orc_triangle_distance_sq(const float* tri9, const float* p) {
    Tri t{v3(tri9[0], tri9[1], tri9[2]), v3(tri9[3], tri9[4], tri9[5]), v3(tri9[6], tri9[7], tri9[8])};
    return tri_distance_sq(t, v3(p[0], p[1], p[2]));
}

int orc_deform_samples(const orc_scene* s, float* coords, uint32_t n, uint32_t stride, int32_t frame) {
    const uint32_t n_extra = s->cfg.n_extra_dims;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        float* c = coords + (uint64_t)i * stride;
        if (c[4] < 0.0f) continue;
        uint32_t f = frame < 0 ? (uint32_t)c[4] : (uint32_t)frame;
        if (f >= s->frames.size() || !s->frames[f].built) continue;
        V3 op, od;
        uint32_t mask = 0;
        if (!warp_one(s, s->frames[f], v3(c[0], c[1], c[2]), v3(c[8], c[9], c[10]), op, od, mask)) continue;
        c[0] = op.x; c[1] = op.y; c[2] = op.z;
        c[8] = od.x; c[9] = od.y; c[10] = od.z;
        if (mask != 7u)
            for (uint32_t m = 0; m < n_extra; ++m) c[stride - n_extra + m] = 1.0f;
    }
    return 0;
}

int orc_build_occupancy(orc_scene* s, int32_t frame, float threshold, uint8_t* out) {
    const MeshFrame* fr = frame_ptr(s, frame);
    if (!fr) return 1;
    const uint32_t G = s->cfg.grid_size;
    const uint32_t n_cells = G * G * G;
    std::vector<uint8_t> bits(n_cells / 8, 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t byte = 0; byte < (int64_t)(n_cells / 8); ++byte) {
        uint8_t v = 0;
        for (uint32_t j = 0; j < 8; ++j) {
            uint32_t idx = (uint32_t)byte * 8 + j;
            uint32_t x = compact3(idx), y = compact3(idx >> 1), z = compact3(idx >> 2);
            V3 pos = v3((((float)x + 0.5f) / (float)G - 0.5f) * 1.0f + 0.5f,
                        (((float)y + 0.5f) / (float)G - 0.5f) * 1.0f + 0.5f,
                        (((float)z + 0.5f) / (float)G - 0.5f) * 1.0f + 0.5f);
            Hit h = bvh_closest(fr->bvh, pos);
            if (h.idx < 0) continue;
            float dist = std::sqrt(h.dist_sq);
            if (dist < threshold && s->masks[fr->bvh.orig[h.idx]] != 0u) v |= (uint8_t)(1u << j);
        }
        bits[byte] = v;
    }
    s->grids[frame < 0 ? 0 : frame + 1] = bits;
    if (out) std::memcpy(out, bits.data(), bits.size());
    return 0;
}

int orc_set_occupancy(orc_scene* s, int32_t slot, const uint8_t* bits) {
    const uint32_t G = s->cfg.grid_size;
    size_t nb = (size_t)G * G * G / 8;
    s->grids[slot < 0 ? 0 : slot + 1].assign(bits, bits + nb);
    return 0;
}

int orc_set_params(orc_scene* s, const float* params, uint64_t n) {
    if (n != s->layout.n_params) return 1;
    s->params.assign(params, params + n);
    return 0;
}

int orc_encode(const orc_scene* s, const float* pos, uint32_t n, float* out) {
    const uint32_t w = s->cfg.n_levels * s->cfg.n_features_per_level;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) encode_one(s, v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), out + i * w);
    return 0;
}

// corner indices of every level (bit-exact KAT surface): idx_out [n][L][8]
int orc_encode_indices(const orc_scene* s, const float* pos, uint32_t n, uint32_t* idx_out) {
    const insta_layout& L = s->layout;
    const uint32_t nl = s->cfg.n_levels;
    for (uint32_t i = 0; i < n; ++i) {
        const float xs[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        for (uint32_t l = 0; l < nl; ++l) {
            uint32_t g[3];
            for (int d = 0; d < 3; ++d) g[d] = (uint32_t)(int32_t)std::floor(xs[d] * L.level_scale[l] + 0.5f);
            for (uint32_t c = 0; c < 8; ++c)
                idx_out[((uint64_t)i * nl + l) * 8 + c] =
                    grid_index(L, l, g[0] + (c & 1u), g[1] + ((c >> 1) & 1u), g[2] + ((c >> 2) & 1u));
        }
    }
    return 0;
}

int orc_encode_backward(const orc_scene* s, const float* pos, uint32_t n, const float* dL, double* grad) {
    const insta_layout& L = s->layout;
    const uint32_t F = s->cfg.n_features_per_level, nl = s->cfg.n_levels;
    for (uint32_t i = 0; i < n; ++i) {
        const float xs[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        for (uint32_t l = 0; l < nl; ++l) {
            float frac[3];
            uint32_t g[3];
            for (int d = 0; d < 3; ++d) {
                float p = xs[d] * L.level_scale[l] + 0.5f;
                float fl = std::floor(p);
                frac[d] = p - fl;
                g[d] = (uint32_t)(int32_t)fl;
            }
            for (uint32_t c = 0; c < 8; ++c) {
                float w = 1.0f;
                for (int d = 0; d < 3; ++d) w = w * (((c >> d) & 1u) ? frac[d] : 1.0f - frac[d]);
                uint32_t idx = grid_index(L, l, g[0] + (c & 1u), g[1] + ((c >> 1) & 1u), g[2] + ((c >> 2) & 1u));
                for (uint32_t f = 0; f < F; ++f)
                    grad[(uint64_t(L.level_offset[l]) + idx) * F + f] += (double)w * (double)dL[((uint64_t)i * nl + l) * F + f];
            }
        }
    }
    return 0;
}

int orc_sh16(const float* dirs_w, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i)
        sh16(dirs_w[3 * i] * 2.0f - 1.0f, dirs_w[3 * i + 1] * 2.0f - 1.0f, dirs_w[3 * i + 2] * 2.0f - 1.0f, out + 16 * i);
    return 0;
}

int orc_network_inference(const orc_scene* s, const float* coords, uint32_t n, uint32_t stride, float* out) {
    const uint32_t n_extra = s->cfg.n_extra_dims;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* c = coords + (uint64_t)i * stride;
        network_one(s, v3(c[0], c[1], c[2]), v3(c[8], c[9], c[10]), n_extra ? c + stride - n_extra : nullptr,
                    out + 4 * i);
    }
    return 0;
}

float orc_ld_random_val(uint32_t index, uint32_t seed, uint32_t dim) { return ld_random_val(index, seed, dim); }
uint32_t orc_sobol(uint32_t index, uint32_t dim) { return sobol_dim(index, dim); }
uint32_t orc_owen_scramble(uint32_t x, uint32_t seed) { return owen_scramble(x, seed); }
uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }
uint32_t orc_morton3_invert(uint32_t v) { return compact3(v); }
void orc_pixel_offset(uint32_t spp, float* out2) { ld_random_pixel_offset(spp, out2); }

// Full render of one frame: init_rays_with_payload_kernel_nerf + advance_pos_nerf
// (testbed_nerf.cu:1854-1963, 634-686), then per ray the sequential body of
// generate_next_nerf_network_inputs + warp + network + composite_kernel_nerf
// (722-795, core.cu:442-526, 797-1009).  Sample-by-sample per ray, which gives
// the same result as the reference's compaction schedule because each ray's
// sample sequence does not depend on it.  rgba [w*h*4], depth [w*h],
// steps [w*h], stats_out[0] = samples composited, [1] = rays alive after init.
int orc_render(const orc_scene* s, const insta_camera* cam, uint32_t width, uint32_t height, int32_t frame,
               int32_t occupancy, uint32_t spp, const uint32_t* pixel_subset, uint32_t n_subset, float* rgba,
               float* depth, uint32_t* steps, uint64_t* stats_out) {
    const MeshFrame* fr = frame_ptr(s, frame);
    if (!fr || s->params.empty()) return 1;
    const std::vector<uint8_t>& grid = s->grids[occupancy < 0 ? 0 : occupancy + 1];
    if (grid.empty()) return 2;
    const insta_config& c = s->cfg;
    const uint32_t G = c.grid_size;
    const float DT = c.dt;
    const V3 lo = s->render_lo, hi = s->render_hi;
    const float* M = cam->matrix;
    const V3 cam_pos = v3(M[3], M[7], M[11]);
    const V3 cam_fwd = v3(M[2], M[6], M[10]);
    float offset[2];
    ld_random_pixel_offset(cam->snap_to_pixel_centers ? 0u : spp, offset);
    const uint32_t n_pix = pixel_subset ? n_subset : width * height;
    uint64_t total_samples = 0, alive0 = 0;
    int nt = 1;
#ifdef _OPENMP
    nt = s->n_threads > 0 ? s->n_threads : omp_get_max_threads();
#endif
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total_samples, alive0) num_threads(nt)
    for (int64_t q = 0; q < (int64_t)n_pix; ++q) {
        const uint32_t i = pixel_subset ? pixel_subset[q] : (uint32_t)q;
        const uint32_t x = i % width, y = i / width;
        float* out = rgba + 4 * (uint64_t)q;
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        if (depth) depth[q] = 1e10f;
        if (steps) steps[q] = 0;
        // pixel_to_ray, pinhole (common_device.cuh:260-315)
        float u = ((float)x + offset[0]) / (float)width;
        float v = ((float)y + offset[1]) / (float)height;
        V3 d0 = v3((u - cam->screen_center[0]) * (float)width / cam->focal_px[0],
                   (v - cam->screen_center[1]) * (float)height / cam->focal_px[1], 1.0f);
        V3 d = v3(M[0] * d0.x + M[1] * d0.y + M[2] * d0.z, M[4] * d0.x + M[5] * d0.y + M[6] * d0.z,
                  M[8] * d0.x + M[9] * d0.y + M[10] * d0.z);
        float dn = norm(d);
        d = v3(d.x / dn, d.y / dn, d.z / dn);
        const V3 o = cam_pos;
        float t = std::fmax(box_ray_tmin(lo, hi, o, d), 0.05f) + 1e-6f;
        if (!box_contains(lo, hi, o + d * t)) continue;
        const V3 idir = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        // advance_pos_nerf: jitter by one step, skip empty space
        t += (ORC_MUTANT == 4 ? 0.0f : ld_random_val(spp, i * 786433u, 0)) * DT;
        bool alive = true;
        while (true) {
            V3 pos = o + d * t;
            if (!box_contains(lo, hi, pos)) {
                alive = false;
                break;
            }
            if (occupied(grid.data(), G, pos)) break;
            t = advance_to_next_voxel(t, DT, pos, d, idir, G);
        }
        if (!alive) continue;
        alive0 += 1;
        float acc[4] = {0, 0, 0, 0};
        float local_depth = 0.0f, max_weight = 0.0f;
        uint32_t n = 0;
        const V3 dir_w = v3((d.x + 1.0f) * 0.5f, (d.y + 1.0f) * 0.5f, (d.z + 1.0f) * 0.5f);
        while (n < c.max_steps) {
            V3 pos;
            bool inside = true;
            while (true) {
                pos = o + d * t;
                if (!box_contains(lo, hi, pos)) {
                    inside = false;
                    break;
                }
                if (occupied(grid.data(), G, pos)) break;
                t = advance_to_next_voxel(t, DT, pos, d, idir, G);
            }
            if (!inside) break;
            // deform (warp) then network on the canonical sample
            V3 cp = pos, cd = dir_w;
            uint32_t mask = 0;
            const bool warped = warp_one(s, *fr, pos, dir_w, cp, cd, mask);
            float net[4];
            network_one(s, cp, cd, sample_extra(s, frame, warped, mask), net);
            // composite_kernel_nerf body
            float T = 1.0f - acc[3];
            float sigma = std::exp(net[3] + c.density_logit_bias);
            float alpha = 1.0f - std::exp(-sigma * DT);
            float weight = alpha * T;
            acc[0] = acc[0] + logistic(net[0]) * weight;
            acc[1] = acc[1] + logistic(net[1]) * weight;
            acc[2] = acc[2] + logistic(net[2]) * weight;
            acc[3] = acc[3] + weight;
            if (weight > max_weight) {
                max_weight = weight;
                local_depth = dot(cam_fwd, pos - cam_pos);
            }
            ++n;
            t += DT;
            if (acc[3] > 1.0f - c.min_transmittance) {
                float a = acc[3];
                acc[0] /= a; acc[1] /= a; acc[2] /= a; acc[3] /= a;
                break;
            }
        }
        total_samples += n;
        if (steps) steps[q] = n;
        if (acc[3] > 0.001f) {
            out[0] = acc[0]; out[1] = acc[1]; out[2] = acc[2]; out[3] = acc[3];
        }
        if (depth && acc[3] > 0.2f) depth[q] = local_depth;
    }
    if (stats_out) {
        stats_out[0] = total_samples;
        stats_out[1] = alive0;
    }
    return 0;
}

// Render-buffer stages of Testbed::render_frame for one sample index:
// shade_kernel_nerf into a cleared frame buffer (testbed_nerf.cu:1794-1827),
// accumulate_kernel in the linear colour space (render_buffer.cu:236-266,
// 560-578) and tonemap_kernel (render_buffer.cu:264-350, 493-522), whose
// result is also written back into the accumulation buffer.
namespace {
float srgb_to_linear_f(float c) { return c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f); }
float linear_to_srgb_f(float c) { return c < 0.0031308f ? 12.92f * c : 1.055f * std::pow(c, 0.41666f) - 0.055f; }
}  // namespace

int orc_frame_buffer(const float* rgba, const float* depth, uint32_t n, uint32_t spp, const insta_frame_config* fc,
                     float* accum, float* out, float* depth_out) {
    if (!fc || fc->tonemap_curve > INSTA_TONEMAP_REINHARD) return 1;
    const float s = (float)spp;
    const float acc_w = ORC_MUTANT == 6 ? s : s + 1.0f;
    const float expo = std::pow(2.0f, fc->exposure);
    float bg[4];
    for (int c = 0; c < 3; ++c) bg[c] = srgb_to_linear_f(fc->background[c]);
    bg[3] = fc->background[3];
    float k[6] = {0, 0, 0, 0, 0, 0};
    if (fc->tonemap_curve == INSTA_TONEMAP_ACES) {
        k[0] = 0.6f * 0.6f * 2.51f; k[1] = 0.6f * 0.03f; k[2] = 0.0f;
        k[3] = 0.6f * 0.6f * 2.43f; k[4] = 0.6f * 0.59f; k[5] = 0.14f;
    } else if (fc->tonemap_curve == INSTA_TONEMAP_HABLE) {
        const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f, W = 11.2f;
        float k0 = A * F - A * E, k1 = C * B * F - B * E, k2 = 0.0f, k3 = A * F, k4 = B * F, k5 = D * F * F;
        const float nom = k0 * (W * W) + k1 * W + k2, den = k3 * (W * W) + k4 * W + k5, white = den / nom;
        k[0] = 4.0f * k0 * white; k[1] = 2.0f * k1 * white; k[2] = k2 * white;
        k[3] = 4.0f * k3; k[4] = 2.0f * k4; k[5] = k5;
    }
    for (uint32_t i = 0; i < n; ++i) {
        float f[4] = {rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2], rgba[4 * i + 3]};
        if (!fc->linear_colors)
            for (int c = 0; c < 3; ++c) f[c] = srgb_to_linear_f(f[c]);
        if (depth_out) depth_out[i] = f[3] > 0.2f && depth ? depth[i] : 0.0f;
        float t[4] = {0, 0, 0, 0};
        if (spp != 0)
            for (int c = 0; c < 4; ++c) t[c] = accum[4 * i + c];
        for (int c = 0; c < 4; ++c) t[c] = (t[c] * s + f[c]) / acc_w;
        const float w = (1.0f - t[3]) * bg[3];
        float col[3];
        for (int c = 0; c < 3; ++c) col[c] = t[c] + bg[c] * w;
        t[3] = t[3] + w;
        for (int c = 0; c < 3; ++c) col[c] = col[c] * expo;
        if (fc->tonemap_curve != INSTA_TONEMAP_IDENTITY) {
            for (int c = 0; c < 3; ++c) col[c] = std::fmax(col[c], 0.0f);
            if (fc->tonemap_curve == INSTA_TONEMAP_REINHARD) {
                const float Y = 0.2126f * col[0] + 0.7152f * col[1] + 0.0722f * col[2];
                const float sc = 1.0f / (Y + 1.0f);
                for (int c = 0; c < 3; ++c) col[c] = col[c] * sc;
            } else {
                for (int c = 0; c < 3; ++c) {
                    const float sq = col[c] * col[c];
                    col[c] = (sq * k[0] + k[1] * col[c] + k[2]) / (k[3] * sq + k[4] * col[c] + k[5]);
                }
            }
        }
        if (fc->output_srgb)
            for (int c = 0; c < 3; ++c) col[c] = linear_to_srgb_f(col[c]);
        for (int c = 0; c < 3; ++c) accum[4 * i + c] = out[4 * i + c] = col[c];
        accum[4 * i + 3] = out[4 * i + 3] = t[3];
    }
    return 0;
}

// FP64 reference render: the FP32 march of orc_render (identical samples),
// then deformation, encoding, MLPs and compositing in double.  rgba [n*4].
int orc_render_fp64(const orc_scene* s, const insta_camera* cam, uint32_t width, uint32_t height, int32_t frame,
                    int32_t occupancy, uint32_t spp, const uint32_t* pixel_subset, uint32_t n_subset, double* rgba) {
    const MeshFrame* fr = frame_ptr(s, frame);
    if (!fr || s->params.empty()) return 1;
    const std::vector<uint8_t>& grid = s->grids[occupancy < 0 ? 0 : occupancy + 1];
    if (grid.empty()) return 2;
    const insta_config& c = s->cfg;
    const uint32_t G = c.grid_size;
    const float DT = c.dt;
    const V3 lo = s->render_lo, hi = s->render_hi;
    const float* M = cam->matrix;
    const V3 cam_pos = v3(M[3], M[7], M[11]);
    float offset[2];
    ld_random_pixel_offset(cam->snap_to_pixel_centers ? 0u : spp, offset);
    const uint32_t n_pix = pixel_subset ? n_subset : width * height;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < (int64_t)n_pix; ++q) {
        const uint32_t i = pixel_subset ? pixel_subset[q] : (uint32_t)q;
        const uint32_t x = i % width, y = i / width;
        double* out = rgba + 4 * (uint64_t)q;
        out[0] = out[1] = out[2] = out[3] = 0.0;
        float u = ((float)x + offset[0]) / (float)width;
        float v = ((float)y + offset[1]) / (float)height;
        V3 d0 = v3((u - cam->screen_center[0]) * (float)width / cam->focal_px[0],
                   (v - cam->screen_center[1]) * (float)height / cam->focal_px[1], 1.0f);
        V3 d = v3(M[0] * d0.x + M[1] * d0.y + M[2] * d0.z, M[4] * d0.x + M[5] * d0.y + M[6] * d0.z,
                  M[8] * d0.x + M[9] * d0.y + M[10] * d0.z);
        float dn = norm(d);
        d = v3(d.x / dn, d.y / dn, d.z / dn);
        const V3 o = cam_pos;
        float t = std::fmax(box_ray_tmin(lo, hi, o, d), 0.05f) + 1e-6f;
        if (!box_contains(lo, hi, o + d * t)) continue;
        const V3 idir = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        t += ld_random_val(spp, i * 786433u, 0) * DT;
        bool alive = true;
        while (true) {
            V3 pos = o + d * t;
            if (!box_contains(lo, hi, pos)) {
                alive = false;
                break;
            }
            if (occupied(grid.data(), G, pos)) break;
            t = advance_to_next_voxel(t, DT, pos, d, idir, G);
        }
        if (!alive) continue;
        double acc[4] = {0, 0, 0, 0};
        uint32_t n = 0;
        const D3v dird = {d.x, d.y, d.z};
        while (n < c.max_steps) {
            V3 pos;
            bool inside = true;
            while (true) {
                pos = o + d * t;
                if (!box_contains(lo, hi, pos)) {
                    inside = false;
                    break;
                }
                if (occupied(grid.data(), G, pos)) break;
                t = advance_to_next_voxel(t, DT, pos, d, idir, G);
            }
            if (!inside) break;
            D3v cp = dv(pos), cd = dird;
            uint32_t mask = 0;
            const bool warped = warp_one_d(s, *fr, dv(pos), dird, cp, cd, mask);
            double net[4];
            network_one_d(s, cp, cd, sample_extra(s, frame, warped, mask), net);
            const double T = 1.0 - acc[3];
            const double sigma = std::exp(net[3] + (double)c.density_logit_bias);
            const double alpha = 1.0 - std::exp(-sigma * (double)DT);
            const double weight = alpha * T;
            for (int k = 0; k < 3; ++k) acc[k] += weight / (1.0 + std::exp(-net[k]));
            acc[3] += weight;
            ++n;
            t += DT;
            if (acc[3] > 1.0 - (double)c.min_transmittance) {
                const double a = acc[3];
                for (int k = 0; k < 4; ++k) acc[k] /= a;
                break;
            }
        }
        if (acc[3] > 0.001)
            for (int k = 0; k < 4; ++k) out[k] = acc[k];
    }
    return 0;
}

}  // extern "C"

// ===========================================================================
// Training step (Testbed::train_nerf_step, instant-ngp/src/testbed_nerf.cu:
// 3176-3400): the FP32 ray generation, march and deformation of the device
// kernels (identical canonical samples), then network, loss and the backward
// pass in FP64.  Gradients are those of the objective
//   L = (1 / n_rays) * sum_rays [ s_r * sum_channels Huber_0.1(rgb - target) / 5
//                                 + l_r * |D_r - D*_r| ],
// i.e. the per-channel gradient of compute_loss_kernel_train_nerf
// (1305-1643) without the loss scale; the reported loss is the per-ray
// channel mean of the colour term without s_r, as the reference reports it.
// With segmentation masks and depth images (1476-1497): s_r is the seg
// weight for mask ids 120, 130, 140, else 1; l_r is the depth weight, 0 for
// mask ids <= 90 or > 250 or texel alpha < 0.001 (only where masks are set),
// and 0 where the target D*_r = |unnormalised direction| * depth texel is not
// positive; D_r = sum_j w_j |x_j - o| over the composited samples (posed
// positions x_j, ray origin o), L1 (the default depth_loss_type).
// ===========================================================================
namespace {

// PCG32 (O'Neill 2014; pcg32.h by W. Jakob, instant-ngp's default_rng_t)
struct Pcg32 {
    uint64_t state, inc;
    static Pcg32 seeded(uint64_t initstate, uint64_t initseq = 0xda3e39cb94b95bdbull) {
        Pcg32 r{0u, (initseq << 1u) | 1u};
        r.next_uint();
        r.state += initstate;
        r.next_uint();
        return r;
    }
    uint32_t next_uint() {
        const uint64_t old = state;
        state = old * 0x5851f42d4c957f2dull + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    float next_float() {
        const uint32_t u = (next_uint() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &u, 4);
        return f - 1.0f;
    }
    void advance(uint64_t delta) {
        uint64_t cur_mult = 0x5851f42d4c957f2dull, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
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

constexpr uint32_t kRaySeedStride = 8;  // N_MAX_RANDOM_SAMPLES_PER_RAY (testbed_nerf.cu:60)
constexpr double kHuberDelta = ORC_MUTANT == 3 ? 0.2 : 0.1;  // testbed_nerf.cu:1299

// sRGB transfer functions (common_device.cuh:31-61), the reference's
// constants (its encoding exponent is 0.41666, not 1/2.4)
inline double srgb_to_linear_d(double c) { return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4); }
inline double linear_to_srgb_d(double c) { return c < 0.0031308 ? 12.92 * c : 1.055 * std::pow(c, 0.41666) - 0.055; }

struct TrainRay {
    uint32_t src, img, pix;
    V3 o, d;
    float dn;  // length of the unnormalised direction (camera z component 1)
    std::vector<V3> samples;
    float bg[3];
};

// generate_training_samples_nerf for ray i (testbed_nerf.cu:1137-1285)
bool train_ray(const orc_scene* s, uint32_t i, uint32_t n_rays, uint64_t rays_total, const Pcg32& rng0,
               TrainRay& r) {
    const uint32_t n_img = s->tr_n, W = s->tr_w, H = s->tr_h;
    r.src = i;
    r.img = (uint32_t)((((uint64_t)i + rays_total) * n_img / n_rays) % n_img);
    Pcg32 rng = rng0;
    rng.advance((uint64_t)i * kRaySeedStride);
    const float x = rng.next_float();
    const float y = rng.next_float();
    (void)rng.next_float();  // motion-blur time
    const float jitter = rng.next_float();
    const uint32_t px = std::min((uint32_t)(x * (float)W), W - 1u);
    const uint32_t py = std::min((uint32_t)(y * (float)H), H - 1u);
    r.pix = px + py * W;
    // background: draws 2..4 of the same stream (testbed_nerf.cu:1408-1419)
    Pcg32 rb = rng0;
    rb.advance((uint64_t)i * kRaySeedStride);
    (void)rb.next_float();
    (void)rb.next_float();
    for (int c = 0; c < 3; ++c) r.bg[c] = rb.next_float();
    const insta_camera& cam = s->tr_cams[r.img];
    const float* M = cam.matrix;
    V3 d0 = v3((x - cam.screen_center[0]) * (float)W / cam.focal_px[0],
               (y - cam.screen_center[1]) * (float)H / cam.focal_px[1], 1.0f);
    V3 d = v3(M[0] * d0.x + M[1] * d0.y + M[2] * d0.z, M[4] * d0.x + M[5] * d0.y + M[6] * d0.z,
              M[8] * d0.x + M[9] * d0.y + M[10] * d0.z);
    const float dn = norm(d);
    d = v3(d.x / dn, d.y / dn, d.z / dn);
    const V3 o = v3(M[3], M[7], M[11]);
    r.o = o;
    r.d = d;
    r.dn = dn;
    r.samples.clear();
    const float tb = box_ray_tmin(s->render_lo, s->render_hi, o, d);
    if (tb == FLT_MAX) return false;
    const std::vector<uint8_t>& grid = s->grids[(s->tr_occ_slot >= -1 ? s->tr_occ_slot : s->tr_frames[r.img]) + 1];
    const uint32_t G = s->cfg.grid_size;
    const float DT = s->cfg.dt;
    const V3 idir = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float t = std::fmax(tb, 0.0f) + DT * jitter;
    while (r.samples.size() < s->cfg.max_steps) {
        V3 pos;
        bool inside = true;
        while (true) {
            pos = o + d * t;
            if (!box_contains(s->render_lo, s->render_hi, pos)) {
                inside = false;
                break;
            }
            if (occupied(grid.data(), G, pos)) break;
            t = advance_to_next_voxel(t, DT, pos, d, idir, G);
        }
        if (!inside) break;
        r.samples.push_back(pos);
        t += DT;
    }
    return !r.samples.empty();
}

struct SampleAct {
    double enc[64], h1[64], cin[32], h2[64], h3[64], out[4];
    D3v pos, dir;
};

// forward of the supported shape (1 density / 2 colour hidden layers, width
// 64, F = 2) in FP64, keeping the activations; the density input is the
// encoding followed by the sample's extra dims
void forward_d(const orc_scene* s, D3v pos, D3v dir, const float* extra, SampleAct& a) {
    const uint32_t E = s->cfg.n_levels * 2, DIN = E + s->cfg.n_extra_dims;
    a.pos = pos;
    a.dir = dir;
    encode_one_d(s, pos, a.enc);
    for (uint32_t e = E; e < DIN; ++e) a.enc[e] = extra[e - E];
    const float* p = s->params.data();
    dense_d(p, a.enc, DIN, 64, true, a.h1);
    p += 64 * DIN;
    dense_d(p, a.h1, 64, 16, false, a.cin);
    p += 16 * 64;
    // SH of the direction (the closed forms of network_one_d)
    const double x = dir.x, y = dir.y, z = dir.z;
    double* o = a.cin + 16;
    const double pi = 3.14159265358979323846;
    const double c0 = 0.5 / std::sqrt(pi), c1 = std::sqrt(3.0 / (4 * pi)), c2 = 0.5 * std::sqrt(15.0 / pi),
                 c3 = 0.25 * std::sqrt(5.0 / pi), c4 = 0.25 * std::sqrt(15.0 / pi), c5 = 0.25 * std::sqrt(35.0 / (2 * pi)),
                 c6 = 0.5 * std::sqrt(105.0 / pi), c7 = 0.25 * std::sqrt(21.0 / (2 * pi)), c8 = 0.25 * std::sqrt(7.0 / pi),
                 c9 = 0.25 * std::sqrt(105.0 / pi);
    o[0] = c0; o[1] = -c1 * y; o[2] = c1 * z; o[3] = -c1 * x;
    o[4] = c2 * x * y; o[5] = -c2 * y * z; o[6] = c3 * (3 * z * z - 1); o[7] = -c2 * x * z;
    o[8] = c4 * (x * x - y * y); o[9] = c5 * y * (3 * x * x - y * y) * -1.0; o[10] = c6 * x * y * z;
    o[11] = c7 * y * (1 - 5 * z * z); o[12] = c8 * z * (5 * z * z - 3); o[13] = c7 * x * (1 - 5 * z * z);
    o[14] = c9 * z * (x * x - y * y); o[15] = c5 * x * (-x * x + 3 * y * y);
    dense_d(p, a.cin, 32, 64, true, a.h2);
    p += 64 * 32;
    dense_d(p, a.h2, 64, 64, true, a.h3);
    p += 64 * 64;
    dense_d(p, a.h3, 64, 3, false, a.out);
    a.out[3] = a.cin[0];
}

// backward of one sample: accumulates dL/dparams into g (FP64); with
// dinput (6 values) also the input gradients of NerfNetwork::backward_impl:
// the position through each level's trilinear weights, the direction (input
// x in [0,1], d = 2x - 1) through the SH polynomials (nerf_network.h:318-380)
void backward_d(const orc_scene* s, const SampleAct& a, const double dout[4], double* g, double* dinput = nullptr) {
    const uint32_t E = s->cfg.n_levels * 2, DIN = E + s->cfg.n_extra_dims;
    const uint64_t W0 = 0, W1 = W0 + 64 * DIN, W2 = W1 + 16 * 64, W3 = W2 + 64 * 32, W4 = W3 + 64 * 64;
    const float* P = s->params.data();
    double dh3[64], dh2[64], dd[16], dh1[64], denc[64];
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 64; ++k) g[W4 + c * 64 + k] += dout[c] * a.h3[k];
    for (int k = 0; k < 64; ++k) {
        double acc = 0.0;
        for (int c = 0; c < 3; ++c) acc += (double)P[W4 + c * 64 + k] * dout[c];
        dh3[k] = a.h3[k] > 0.0 ? acc : 0.0;
    }
    for (int m = 0; m < 64; ++m)
        for (int k = 0; k < 64; ++k) g[W3 + m * 64 + k] += dh3[m] * a.h2[k];
    for (int k = 0; k < 64; ++k) {
        double acc = 0.0;
        for (int m = 0; m < 64; ++m) acc += (double)P[W3 + m * 64 + k] * dh3[m];
        dh2[k] = a.h2[k] > 0.0 ? acc : 0.0;
    }
    for (int m = 0; m < 64; ++m)
        for (int q = 0; q < 32; ++q) g[W2 + m * 32 + q] += dh2[m] * a.cin[q];
    for (int q = 0; q < 16; ++q) {
        double acc = 0.0;
        for (int m = 0; m < 64; ++m) acc += (double)P[W2 + m * 32 + q] * dh2[m];
        dd[q] = acc;
    }
    dd[0] += dout[3];
    for (int q = 0; q < 16; ++q)
        for (int k = 0; k < 64; ++k) g[W1 + q * 64 + k] += dd[q] * a.h1[k];
    for (int k = 0; k < 64; ++k) {
        double acc = 0.0;
        for (int q = 0; q < 16; ++q) acc += (double)P[W1 + q * 64 + k] * dd[q];
        dh1[k] = a.h1[k] > 0.0 ? acc : 0.0;
    }
    for (int k = 0; k < 64; ++k)
        for (uint32_t e = 0; e < DIN; ++e) g[W0 + k * DIN + e] += dh1[k] * a.enc[e];
    for (uint32_t e = 0; e < E; ++e) {  // the extra dims get no gradient (no latent-code optimisation)
        double acc = 0.0;
        for (int k = 0; k < 64; ++k) acc += (double)P[W0 + k * DIN + e] * dh1[k];
        denc[e] = acc;
    }
    // hash-grid scatter with the FP64 trilinear weights of encode_one_d
    const insta_layout& L = s->layout;
    double* gt = g + L.n_mlp_params;
    const double xs[3] = {a.pos.x, a.pos.y, a.pos.z};
    for (uint32_t l = 0; l < s->cfg.n_levels; ++l) {
        double frac[3];
        uint32_t gi[3];
        for (int d = 0; d < 3; ++d) {
            double pp = xs[d] * (double)L.level_scale[l] + 0.5;
            double fl = std::floor(pp);
            frac[d] = pp - fl;
            gi[d] = (uint32_t)(int32_t)fl;
        }
        for (uint32_t c = 0; c < 8; ++c) {
            double w = 1.0;
            for (int d = 0; d < 3; ++d) w *= ((c >> d) & 1u) ? frac[d] : 1.0 - frac[d];
            uint32_t idx = grid_index(L, l, gi[0] + (c & 1u), gi[1] + ((c >> 1) & 1u), gi[2] + ((c >> 2) & 1u));
            const uint64_t e = (uint64_t(L.level_offset[l]) + idx) * 2;
            gt[e] += w * denc[2 * l];
            gt[e + 1] += w * denc[2 * l + 1];
        }
    }
    if (!dinput) return;
    // d/dx of the trilinear interpolation: per corner, the derivative of its
    // weight along each axis (scale_l times the other two factors)
    const float* table = P + L.n_mlp_params;
    double gx[3] = {0, 0, 0};
    for (uint32_t l = 0; l < s->cfg.n_levels; ++l) {
        const double sc = (double)L.level_scale[l];
        double frac[3];
        uint32_t gi[3];
        for (int d = 0; d < 3; ++d) {
            double pp = xs[d] * sc + 0.5;
            double fl = std::floor(pp);
            frac[d] = pp - fl;
            gi[d] = (uint32_t)(int32_t)fl;
        }
        for (uint32_t c = 0; c < 8; ++c) {
            double f[3];
            for (int d = 0; d < 3; ++d) f[d] = ((c >> d) & 1u) ? frac[d] : 1.0 - frac[d];
            uint32_t idx = grid_index(L, l, gi[0] + (c & 1u), gi[1] + ((c >> 1) & 1u), gi[2] + ((c >> 2) & 1u));
            const uint64_t e = (uint64_t(L.level_offset[l]) + idx) * 2;
            const double sfe = denc[2 * l] * (double)table[e] + denc[2 * l + 1] * (double)table[e + 1];
            for (int d = 0; d < 3; ++d) {
                const double dw = (((c >> d) & 1u) ? sc : -sc) * f[(d + 1) % 3] * f[(d + 2) % 3];
                gx[d] += sfe * dw;
            }
        }
    }
    // dL/dSH = W2[:, 16:32]^T dH2, through the derivatives of the 16 closed forms
    double dsh[16];
    for (int q = 0; q < 16; ++q) {
        double acc = 0.0;
        for (int m = 0; m < 64; ++m) acc += (double)P[W2 + m * 32 + 16 + q] * dh2[m];
        dsh[q] = acc;
    }
    const double x = a.dir.x, y = a.dir.y, z = a.dir.z, pi = 3.14159265358979323846;
    const double c1 = std::sqrt(3.0 / (4 * pi)), c2 = 0.5 * std::sqrt(15.0 / pi), c3 = 0.25 * std::sqrt(5.0 / pi),
                 c4 = 0.25 * std::sqrt(15.0 / pi), c5 = 0.25 * std::sqrt(35.0 / (2 * pi)), c6 = 0.5 * std::sqrt(105.0 / pi),
                 c7 = 0.25 * std::sqrt(21.0 / (2 * pi)), c8 = 0.25 * std::sqrt(7.0 / pi), c9 = 0.25 * std::sqrt(105.0 / pi);
    // rows: d sh_q / d(x, y, z)
    const double J[16][3] = {{0, 0, 0},
                             {0, -c1, 0},
                             {0, 0, c1},
                             {-c1, 0, 0},
                             {c2 * y, c2 * x, 0},
                             {0, -c2 * z, -c2 * y},
                             {0, 0, 6 * c3 * z},
                             {-c2 * z, 0, -c2 * x},
                             {2 * c4 * x, -2 * c4 * y, 0},
                             {-6 * c5 * x * y, c5 * (3 * y * y - 3 * x * x), 0},
                             {c6 * y * z, c6 * x * z, c6 * x * y},
                             {0, c7 * (1 - 5 * z * z), -10 * c7 * y * z},
                             {0, 0, c8 * (15 * z * z - 3)},
                             {c7 * (1 - 5 * z * z), 0, -10 * c7 * x * z},
                             {2 * c9 * x * z, -2 * c9 * y * z, c9 * (x * x - y * y)},
                             {c5 * (3 * y * y - 3 * x * x), 6 * c5 * x * y, 0}};
    double gd[3] = {0, 0, 0};
    for (int q = 0; q < 16; ++q)
        for (int d = 0; d < 3; ++d) gd[d] += dsh[q] * J[q][d];
    for (int d = 0; d < 3; ++d) {
        dinput[d] = gx[d];
        dinput[3 + d] = 2.0 * gd[d];
    }
}

}  // namespace

extern "C" {

// NerfNetwork plugin in FP64 (nerf_network.h:238-414): forward on
// NerfCoordinate samples given in double (stride floats each; extra dims the
// last n_extra_dims), out [n*4] = rgb logits and density logit; with dL_dout,
// grads [n_params] (overwritten) and dinput [n*6] (position, then the
// direction input in [0,1]).
int orc_network_plugin_d(const orc_scene* s, const double* coords, uint32_t n, uint32_t stride, const double* dL_dout,
                         double* out, double* grads, double* dinput) {
    const insta_config& c = s->cfg;
    if (s->params.empty() || c.n_features_per_level != 2 || c.n_neurons != 64 || c.density_hidden_layers != 1 ||
        c.rgb_hidden_layers != 2 || c.density_out != 16 || c.n_levels * 2 + c.n_extra_dims > 64 ||
        stride != 11 + c.n_extra_dims)
        return 1;
    const uint64_t np = s->layout.n_params;
    if (grads) std::fill(grads, grads + np, 0.0);
    std::vector<double> g(grads ? np : 0, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        const double* co = coords + (size_t)i * stride;
        float extra[64];
        for (uint32_t e = 0; e < c.n_extra_dims; ++e) extra[e] = (float)co[stride - c.n_extra_dims + e];
        SampleAct a;
        forward_d(s, {co[0], co[1], co[2]}, {co[8] * 2.0 - 1.0, co[9] * 2.0 - 1.0, co[10] * 2.0 - 1.0}, extra, a);
        if (out)
            for (int k = 0; k < 4; ++k) out[4 * (size_t)i + k] = a.out[k];
        if (dL_dout && (grads || dinput)) {
            const double* d = dL_dout + 4 * (size_t)i;
            std::vector<double> tmp;
            double* gg = grads;
            if (!gg) {
                tmp.assign(np, 0.0);
                gg = tmp.data();
            }
            backward_d(s, a, d, gg, dinput ? dinput + 6 * (size_t)i : nullptr);
        }
    }
    return 0;
}

// NerfNetwork::density in FP64: the density MLP's 16 outputs, extra dims 1.0
int orc_network_density_d(const orc_scene* s, const double* pos, uint32_t n, double* out) {
    const insta_config& c = s->cfg;
    if (s->params.empty() || c.density_hidden_layers != 1 || c.density_out != 16) return 1;
    const uint32_t E = c.n_levels * c.n_features_per_level, DIN = E + c.n_extra_dims, W = c.n_neurons;
    for (uint32_t i = 0; i < n; ++i) {
        double din[INSTA_MAX_LEVELS * 8 + 64] = {0}, h[256];
        encode_one_d(s, {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, din);
        for (uint32_t e = E; e < DIN; ++e) din[e] = 1.0;
        dense_d(s->params.data(), din, DIN, W, true, h);
        dense_d(s->params.data() + (uint64_t)W * DIN, h, W, 16, false, out + 16 * (size_t)i);
    }
    return 0;
}

// NerfNetwork::density in FP32 (nerf_network.h:385-414): the density MLP's
// 16 outputs, extra dims 1.0, each layer summed in input order
int orc_network_density(const orc_scene* s, const float* pos, uint32_t n, float* out) {
    const insta_config& c = s->cfg;
    if (s->params.empty() || c.density_out != 16) return 1;
    const uint32_t E = c.n_levels * c.n_features_per_level, DIN = E + c.n_extra_dims, W = c.n_neurons;
    for (uint32_t i = 0; i < n; ++i) {
        float din[INSTA_MAX_LEVELS * 8 + 64], h0[256], h1[256];
        encode_one(s, {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, din);
        for (uint32_t e = E; e < DIN; ++e) din[e] = 1.0f;
        const float* p = s->params.data();
        dense(p, din, DIN, W, true, h0);
        p += (uint64_t)W * DIN;
        for (uint32_t k = 1; k < c.density_hidden_layers; ++k) {
            dense(p, h0, W, W, true, h1);
            std::memcpy(h0, h1, W * sizeof(float));
            p += (uint64_t)W * W;
        }
        dense(p, h0, W, 16, false, out + 16 * (size_t)i);
    }
    return 0;
}

// the grid training rays march: slot >= -1 for all rays, -2 each image's frame
int orc_set_training_occupancy(orc_scene* s, int32_t slot) {
    if (slot < -2 || slot + 1 >= (int32_t)s->grids.size()) return 1;
    s->tr_occ_slot = slot;
    return 0;
}

int orc_set_training_set(orc_scene* s, uint32_t n_images, uint32_t width, uint32_t height, const float* rgba,
                         const insta_camera* cams, const int32_t* frames) {
    for (uint32_t i = 0; i < n_images; ++i)
        if (frames[i] < 0 || (uint32_t)frames[i] >= s->frames.size() || !s->frames[frames[i]].built) return 1;
    s->tr_n = n_images;
    s->tr_w = width;
    s->tr_h = height;
    s->tr_images.assign(rgba, rgba + (size_t)n_images * width * height * 4);
    s->tr_cams.assign(cams, cams + n_images);
    s->tr_frames.assign(frames, frames + n_images);
    s->tr_seg.clear();
    s->tr_depth.clear();
    return 0;
}

// Segmentation masks (mask id per pixel: the red channel of the reference's
// RGBA mask image) and depth images (depth along the camera axis in scene
// units, 0 = none) of the training set, either nullable; weights are
// seg_mask_supervision_lambda and depth_supervision_lambda (testbed.h:622-623,
// src/core.cu:187, 208)
int orc_set_training_supervision(orc_scene* s, const uint8_t* seg, const float* depth, double seg_weight,
                                 double depth_weight) {
    if (s->tr_n == 0) return 1;
    const size_t n = (size_t)s->tr_n * s->tr_w * s->tr_h;
    if (seg) s->tr_seg.assign(seg, seg + n); else s->tr_seg.clear();
    if (depth) s->tr_depth.assign(depth, depth + n); else s->tr_depth.clear();
    s->tr_seg_weight = seg_weight;
    s->tr_depth_weight = depth_weight;
    return 0;
}

void orc_pcg32_draws(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out) {
    Pcg32 r = Pcg32::seeded(initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next_uint();
}

uint32_t orc_pcg32_next(uint64_t initstate, uint64_t initseq, uint64_t skip) {
    Pcg32 r = Pcg32::seeded(initstate, initseq);
    r.advance(skip);
    return r.next_uint();
}

// One training batch: rays i < n_rays of the stream seeded by `seed` and
// advanced by rays_total * 8 (the device's per-step base).  Every ray with
// samples is used (choose max_steps * n_rays within the device's sample cap
// when comparing).  grads [n_params] (nullable) receives dL/dparams in FP64;
// loss[0] = reported loss, loss[1] = the objective L above;
// counts = {rays used, samples, samples composited}.
int orc_train_grads(const orc_scene* s, uint32_t n_rays, uint32_t seed, uint64_t rays_total, double* grads,
                    double* loss, uint32_t* counts) {
    const insta_config& c = s->cfg;
    if (s->tr_n == 0 || s->params.empty() || c.n_features_per_level != 2 || c.n_neurons != 64 ||
        c.density_hidden_layers != 1 || c.rgb_hidden_layers != 2 || c.density_out != 16 ||
        c.n_levels * 2 + c.n_extra_dims > 64)
        return 1;
    // the grid(s) the rays march must exist
    if (s->tr_occ_slot >= -1) {
        if (s->grids[s->tr_occ_slot + 1].empty()) return 1;
    } else {
        for (const int32_t f : s->tr_frames)
            if (s->grids[f + 1].empty()) return 1;
    }
    Pcg32 rng0 = Pcg32::seeded(seed);
    rng0.advance(rays_total * kRaySeedStride);
    const uint64_t np = s->layout.n_params;
    const double dt = (double)c.dt, bias = (double)c.density_logit_bias, inv_rays = 1.0 / (double)n_rays;
    int nt = 1;
#ifdef _OPENMP
    nt = s->n_threads > 0 ? s->n_threads : omp_get_max_threads();
#endif
    std::vector<std::vector<double>> tg(grads ? nt : 0, std::vector<double>(grads ? np : 0, 0.0));
    double rep = 0.0, obj = 0.0;
    uint64_t rays_used = 0, n_samples = 0, n_used = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt) reduction(+ : rep, obj, rays_used, n_samples, n_used)
    for (int64_t i = 0; i < (int64_t)n_rays; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        TrainRay r;
        if (!train_ray(s, (uint32_t)i, n_rays, rays_total, rng0, r)) continue;
        rays_used += 1;
        const uint32_t cnt = (uint32_t)r.samples.size();
        n_samples += cnt;
        const MeshFrame& fr = s->frames[s->tr_frames[r.img]];
        const V3 dir_w = v3((r.d.x + 1.0f) * 0.5f, (r.d.y + 1.0f) * 0.5f, (r.d.z + 1.0f) * 0.5f);
        std::vector<SampleAct> act(cnt);
        for (uint32_t j = 0; j < cnt; ++j) {
            // FP32 deformation (the canonical samples the device path sees:
            // the nearest triangle is a discontinuous choice, so an FP64
            // deformation would move near-tie samples to another triangle's
            // frame); network, loss and backward in FP64 from there
            V3 cp = r.samples[j], cd = dir_w;
            uint32_t mask = 0;
            const bool warped = warp_one(s, fr, r.samples[j], dir_w, cp, cd, mask);
            const D3v dird = {(double)cd.x * 2.0 - 1.0, (double)cd.y * 2.0 - 1.0, (double)cd.z * 2.0 - 1.0};
            forward_d(s, dv(cp), dird, sample_extra(s, s->tr_frames[r.img], warped, mask), act[j]);
        }
        const float* tex = &s->tr_images[((size_t)r.img * s->tr_w * s->tr_h + r.pix) * 4];
        double target[3], bg[3];
        for (int k = 0; k < 3; ++k) {
            // compute_loss_kernel_train_nerf, linear colour space, sRGB
            // training (testbed_nerf.cu:1419-1443): the random background is
            // taken to linear, blended behind the premultiplied linear texel
            // (read_rgba of an 8-bit image, common_device.cuh:658-671), and
            // both the target and the background behind the prediction go
            // back to sRGB
            const double bg_lin = srgb_to_linear_d(r.bg[k]);
            const double a = tex[3];
            if (ORC_MUTANT == 5) {
                bg[k] = r.bg[k];
                target[k] = (double)tex[k] + (1.0 - a) * bg[k];
            } else {
                bg[k] = linear_to_srgb_d(bg_lin);
                target[k] = linear_to_srgb_d(srgb_to_linear_d(tex[k]) * a + (1.0 - a) * bg_lin);
            }
        }
        double T = 1.0, acc[3] = {0, 0, 0};
        uint32_t used = 0;
        for (; used < cnt; ++used) {
            if (T < 1e-4) break;
            const double* o = act[used].out;
            const double sigma = std::exp(o[3] + bias);
            const double alpha = 1.0 - std::exp(-sigma * dt);
            const double w = alpha * T;
            for (int k = 0; k < 3; ++k) acc[k] += w / (1.0 + std::exp(-o[k]));
            T *= 1.0 - alpha;
        }
        n_used += used;
        if (used == cnt)
            for (int k = 0; k < 3; ++k) acc[k] += T * bg[k];
        // segmentation and depth supervision (testbed_nerf.cu:1476-1497)
        const size_t texel = (size_t)r.img * s->tr_w * s->tr_h + r.pix;
        double seg_w = 1.0, lam = s->tr_depth_weight;
        if (!s->tr_seg.empty()) {
            const int id = s->tr_seg[texel];
            if (id == 120 || id == 130 || id == 140) seg_w = s->tr_seg_weight;
            if (id <= 90 || id > 250) lam = 0.0;
            if (tex[3] < 0.001f) lam = 0.0;
        }
        const double dscale = ORC_MUTANT == 7 ? 1.0 : (double)r.dn;
        const double target_depth = dscale * ((lam > 0.0 && !s->tr_depth.empty()) ? (double)s->tr_depth[texel] : -1.0);
        // per-sample depth |x_j - o| and the composited depth
        std::vector<double> depth(target_depth > 0.0 ? used : 0);
        double depth_ray = 0.0;
        if (target_depth > 0.0) {
            double Td = 1.0;
            for (uint32_t j = 0; j < used; ++j) {
                const double dx = (double)r.samples[j].x - r.o.x, dy = (double)r.samples[j].y - r.o.y,
                             dz = (double)r.samples[j].z - r.o.z;
                depth[j] = std::sqrt(dx * dx + dy * dy + dz * dz);
                const double alpha = 1.0 - std::exp(-std::exp(act[j].out[3] + bias) * dt);
                depth_ray += alpha * Td * depth[j];
                Td *= 1.0 - alpha;
            }
        }
        // L1 (testbed_nerf.cu:133-139): gradient copysign(1, D - D*)
        const double dgrad = target_depth > 0.0 ? lam * std::copysign(1.0, depth_ray - target_depth) : 0.0;
        double grad[3], lsum = 0.0;
        for (int k = 0; k < 3; ++k) {
            // huber_loss(target, prediction, 0.1) / 5 (testbed_nerf.cu:141-157, 1299)
            const double diff = acc[k] - target[k], ad = std::fabs(diff);
            const double l = ad > kHuberDelta ? ad - 0.5 * kHuberDelta : (0.5 / kHuberDelta) * diff * diff;
            grad[k] = seg_w * (ad > kHuberDelta ? (diff > 0 ? 1.0 : -1.0) : diff / kHuberDelta) / 5.0;
            lsum += l / 5.0;
        }
        rep += lsum / 3.0 * inv_rays;
        obj += (seg_w * lsum + (target_depth > 0.0 ? lam * std::fabs(depth_ray - target_depth) : 0.0)) * inv_rays;
        if (!grads) continue;
        T = 1.0;
        double acc2[3] = {0, 0, 0}, depth2 = 0.0;
        for (uint32_t j = 0; j < used; ++j) {
            const double* o = act[j].out;
            double sg[3];
            for (int k = 0; k < 3; ++k) sg[k] = 1.0 / (1.0 + std::exp(-o[k]));
            const double sigma = std::exp(o[3] + bias);
            const double alpha = 1.0 - std::exp(-sigma * dt);
            const double w = alpha * T;
            for (int k = 0; k < 3; ++k) acc2[k] += w * sg[k];
            T *= 1.0 - alpha;
            double dout[4], dsig = 0.0;
            for (int k = 0; k < 3; ++k) {
                dout[k] = inv_rays * w * grad[k] * sg[k] * (1.0 - sg[k]);
                dsig += grad[k] * (T * sg[k] - (acc[k] - acc2[k]));
            }
            if (dgrad != 0.0) {
                // d D / d sigma_j = dt (T_{j+1} depth_j - suffix_j)
                depth2 += w * depth[j];
                dsig += dgrad * (T * depth[j] - (ORC_MUTANT == 8 ? 0.0 : depth_ray - depth2));
            }
            dout[3] = inv_rays * sigma * dt * dsig;
            backward_d(s, act[j], dout, tg[tid].data());
        }
    }
    if (grads) {
        for (uint64_t p2 = 0; p2 < np; ++p2) {
            double sum = 0.0;
            for (int t = 0; t < nt; ++t) sum += tg[t][p2];
            grads[p2] = sum;
        }
    }
    if (loss) {
        loss[0] = rep;
        loss[1] = obj;
    }
    if (counts) {
        counts[0] = (uint32_t)rays_used;
        counts[1] = (uint32_t)n_samples;
        counts[2] = (uint32_t)n_used;
    }
    return 0;
}

}  // extern "C"

// ===========================================================================
// warp_random_samples (src/core.cu:528-579) and the density-grid update of
// training (Testbed::update_density_grid_nerf, testbed_nerf.cu:2774-2878 with
// the kernels at 487-577, 585-609), one cascade.
// ===========================================================================
namespace {

// frame of position i: PCG32 advanced by i * (n_frames - 1) draws,
// floor(u * (n_frames - 1)); the last frame is never drawn (reference quirk)
void deform_position(const orc_scene* s, float* p3, uint32_t i, uint32_t n_frames, const Pcg32& rng0) {
    Pcg32 rng = rng0;
    const uint32_t nm = n_frames - 1;
    rng.advance((uint64_t)i * nm);
    const uint32_t frame = (uint32_t)(rng.next_float() * (float)nm);
    const MeshFrame& fr = s->frames[frame];
    const V3 p = v3(p3[0], p3[1], p3[2]);
    Hit h = bvh_closest(fr.bvh, p);
    if (h.idx < 0) return;
    const uint32_t id = fr.bvh.orig[h.idx];
    float W[12], C[12];
    if (s->masks[id] == 2u && !s->cfg.disable_interp) {
        float Wb[12];
        blend_frames(s, fr, id, p, Wb);
        invert_affine_f(Wb, W);
        blend_frames(s, s->canon, id, p, C);
    } else {
        std::memcpy(W, fr.proj_inv[id].m, sizeof(W));
        std::memcpy(C, s->canon.proj[id].m, sizeof(C));
    }
    const V3 c = apply_point(C, apply_point(W, p));
    p3[0] = c.x;
    p3[1] = c.y;
    p3[2] = c.z;
}

// generate_grid_samples_nerf_nonuniform for one sample
void grid_sample(uint32_t i, uint32_t n, uint32_t step, float thresh, uint32_t G, const std::vector<float>& grid,
                 const Pcg32& rng0, float* pos, uint32_t& idx) {
    const uint32_t cells = G * G * G;
    Pcg32 rng = rng0;
    rng.advance((uint64_t)i * 4u);
    (void)rng.next_float();
    for (uint32_t j = 0; j < 10; ++j) {
        idx = ((i + step * n) * 56924617u + j * 19349663u + 96925573u) % cells;
        if (grid[idx] > thresh) break;
    }
    const uint32_t x = compact3(idx), y = compact3(idx >> 1), z = compact3(idx >> 2);
    const float rx = rng.next_float(), ry = rng.next_float(), rz = rng.next_float();
    const float g = (float)G;
    pos[0] = (((float)x + rx) / g - 0.5f) * 1.0f + 0.5f;
    pos[1] = (((float)y + ry) / g - 0.5f) * 1.0f + 0.5f;
    pos[2] = (((float)z + rz) / g - 0.5f) * 1.0f + 0.5f;
}

// density-MLP output 0 (the network's density() path: extra dims all 1.0,
// extract_expressions_random, nerf_network.h:107-127)
float density_logit(const orc_scene* s, V3 x) {
    const insta_config& c = s->cfg;
    const insta_layout& L = s->layout;
    float din[INSTA_MAX_LEVELS * 8 + 64], h[256], out;
    encode_one(s, x, din);
    const uint32_t enc = c.n_levels * c.n_features_per_level;
    for (uint32_t e = 0; e < c.n_extra_dims; ++e) din[enc + e] = 1.0f;
    dense(s->params.data(), din, L.density_in, c.n_neurons, true, h);
    dense(s->params.data() + (uint64_t)c.n_neurons * L.density_in, h, c.n_neurons, 1, false, &out);
    return out;
}

// mark_untrained_density_grid (testbed_nerf.cu:391-431, called at 2795-2806):
// a cell no training camera sees (its centre, widened by the cell's
// half-diagonal, outside every camera's centred frustum of half-extent
// resolution / 2 / focal) is -1 (untrained: never sampled, never occupied);
// a seen cell is 0 if it was untrained, or always when `clear`.  One cascade
// (level 0).
void mark_untrained(orc_scene* s, bool clear) {
    const uint32_t G = s->cfg.grid_size, cells = G * G * G;
    const float radius = 0.5f * 1.73205080757f * 1.0f / (float)G;
    const float half_x = (float)s->tr_w * 0.5f, half_y = (float)s->tr_h * 0.5f;
    for (uint32_t i = 0; i < cells; ++i) {
        const float x = (float)compact3(i), y = (float)compact3(i >> 1), z = (float)compact3(i >> 2);
        const float px = ((x + 0.5f) / (float)G - 0.5f) * 1.0f + 0.5f;
        const float py = ((y + 0.5f) / (float)G - 0.5f) * 1.0f + 0.5f;
        const float pz = ((z + 0.5f) / (float)G - 0.5f) * 1.0f + 0.5f;
        int count = 0;
        for (uint32_t j = 0; j < s->tr_n; ++j) {
            const float* M = s->tr_cams[j].matrix;  // columns right, down, forward, position
            const float lx = px - M[3], ly = py - M[7], lz = pz - M[11];
            const float cx = lx * M[0] + ly * M[4] + lz * M[8];
            const float cy = lx * M[1] + ly * M[5] + lz * M[9];
            const float cz = lx * M[2] + ly * M[6] + lz * M[10];
            if (cz > 0.0f && std::fabs(cx) - radius < cz / s->tr_cams[j].focal_px[0] * half_x &&
                std::fabs(cy) - radius < cz / s->tr_cams[j].focal_px[1] * half_y) {
                count = 1;
                break;
            }
        }
        float& g = s->dg_grid[i];
        if (clear || (g < 0.0f) != (count <= 0)) g = count > 0 ? 0.0f : -1.0f;
    }
}

}  // namespace

extern "C" {

int orc_deform_positions(const orc_scene* s, float* pos, uint32_t n, uint32_t n_frames, uint32_t seed,
                         uint64_t rng_offset) {
    if (n_frames == 0 || n_frames > s->frames.size()) return 1;
    Pcg32 rng = Pcg32::seeded(seed);
    rng.advance(rng_offset);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) deform_position(s, pos + 3 * i, (uint32_t)i, n_frames, rng);
    return 0;
}

// One update; slot -1 = canonical slot.  grid_out [G^3] (nullable), mean_out
// (nullable).  The mean is summed as the device does: blocks of 1024 cells
// reduced by a 256-wide tree, then the block sums by a 256-wide tree.
int orc_update_density_grid(orc_scene* s, int32_t slot, uint32_t n_frames, float decay, uint32_t n_uniform,
                            uint32_t n_nonuniform, uint32_t seed, float* grid_out, float* mean_out) {
    const insta_config& c = s->cfg;
    const uint32_t G = c.grid_size, cells = G * G * G, n = n_uniform + n_nonuniform;
    if (s->params.empty() || n_frames == 0) return 1;
    if (s->dg_step == 0) {
        s->dg_grid.assign(cells, 0.0f);
        Pcg32 r = Pcg32::seeded(seed);
        s->dg_rng_state = r.state;
        s->dg_rng_inc = r.inc;
    }
    // cells no training camera sees, on the first update and whenever the
    // number of training images changed (the reference keys this to training
    // step 0, where its first update runs)
    if (s->tr_n > 0 && (s->dg_step == 0 || s->tr_n != s->dg_marked_n)) {
        mark_untrained(s, s->dg_step == 0);
        s->dg_marked_n = s->tr_n;
    }
    const Pcg32 rng{s->dg_rng_state, s->dg_rng_inc};
    std::vector<float> pos((size_t)n * 3);
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n_uniform; ++i)
        grid_sample(i, n_uniform, s->dg_step, -0.01f, G, s->dg_grid, rng, &pos[3 * (size_t)i], idx[i]);
    for (uint32_t i = 0; i < n_nonuniform; ++i)
        grid_sample(i, n_nonuniform, s->dg_step, 0.01f, G, s->dg_grid, rng, &pos[3 * ((size_t)n_uniform + i)],
                    idx[n_uniform + i]);
    Pcg32 adv = rng;
    adv.advance(1ull << 32);
    s->dg_rng_state = adv.state;
    Pcg32 wr = Pcg32::seeded(seed ^ 0x9e3779b9u);
    wr.advance((uint64_t)s->dg_step << 32);
    std::vector<float> thick(n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        deform_position(s, &pos[3 * i], (uint32_t)i, n_frames, wr);
        const float logit = density_logit(s, v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
        thick[i] = std::exp(logit + c.density_logit_bias) * c.dt;
    }
    std::vector<float> tmp(cells, 0.0f);
    for (uint32_t i = 0; i < n; ++i) tmp[idx[i]] = std::max(tmp[idx[i]], thick[i]);
    for (uint32_t i = 0; i < cells; ++i) {
        const float prev = s->dg_grid[i];
        s->dg_grid[i] = prev < 0.0f ? prev : std::fmax(prev * decay, tmp[i]);
    }
    // mean, in the device's summation order
    const uint32_t nb = (cells + 1023) / 1024;
    const float inv = 1.0f / (float)cells;
    std::vector<float> part(nb);
    for (uint32_t b = 0; b < nb; ++b) {
        float lane[256];
        for (uint32_t t = 0; t < 256; ++t) {
            float acc = 0.0f;
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t i = b * 1024 + k * 256 + t;
                if (i < cells) acc += std::fmax(s->dg_grid[i], 0.0f) * inv;
            }
            lane[t] = acc;
        }
        for (uint32_t o = 128; o > 0; o >>= 1)
            for (uint32_t t = 0; t < o; ++t) lane[t] += lane[t + o];
        part[b] = lane[0];
    }
    float lane[256];
    for (uint32_t t = 0; t < 256; ++t) {
        float acc = 0.0f;
        for (uint32_t i = t; i < nb; i += 256) acc += part[i];
        lane[t] = acc;
    }
    for (uint32_t o = 128; o > 0; o >>= 1)
        for (uint32_t t = 0; t < o; ++t) lane[t] += lane[t + o];
    const float mean = lane[0];
    const float thresh = std::fmin(0.01f, mean);
    std::vector<uint8_t>& bits = s->grids[slot < 0 ? 0 : slot + 1];
    bits.assign(cells / 8, 0);
    for (uint32_t i = 0; i < cells; ++i)
        if (s->dg_grid[i] > thresh) bits[i / 8] |= (uint8_t)(1u << (i % 8));
    if (grid_out) std::memcpy(grid_out, s->dg_grid.data(), (size_t)cells * 4);
    if (mean_out) *mean_out = mean;
    ++s->dg_step;
    return 0;
}

void orc_reset_density_grid(orc_scene* s) { s->dg_step = 0; }

// copy of an occupancy slot's bitfield (slot -1 = canonical); 1 if unset
int orc_get_occupancy(const orc_scene* s, int32_t slot, uint8_t* out) {
    const std::vector<uint8_t>& g = s->grids[slot < 0 ? 0 : slot + 1];
    if (g.empty()) return 1;
    std::memcpy(out, g.data(), g.size());
    return 0;
}

}  // extern "C"
