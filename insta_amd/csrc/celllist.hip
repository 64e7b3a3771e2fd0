// celllist.hip -- per-cell candidate lists for the nearest-triangle query of
// the sample deformation (`warp`, src/core.cu:442-526, which calls
// closest_triangle, bvh.cuh:48-96, once per sample).
//
// The reference answers every query by a BVH traversal.  Here the marched
// samples of a frame all lie in occupied cells of the frame's occupancy grid,
// so the build below gives each occupied cell (split into S^3 sub-cells) the
// short list of triangles that can be nearest to a point inside it; a query
// in such a sub-cell scans that list instead of walking the tree.  The
// invariant that makes the answer identical to the BVH's (and to brute
// force) is stated with the layout in common.hpp.
//
// Build, per slot, all on the device and capturable into a graph (no host
// reads, no memsets, grids sized by the caps):
//   1. words: snapshot of the linear occupancy bits, popcount per 64 cells
//   2. scan:  exclusive prefix of the popcounts (one workgroup)
//   3. cells: linear index of each occupied cell, in rank order
//   4. count: per sub-cell, nearest triangle at the centre (BVH), the bound
//             U^2 from the 8 corners, a bounded BVH range walk counting the
//             triangles whose box is within U; an atomic claims the
//             sub-cell's entries (layout order varies, contents do not)
//   5. fill:  the same walk writes the entries, seed first
//   6. sort:  one wave per list orders the entries after the seed by box
//             distance (rank by counting)
//   7. finish: lists valid unless the occupied cells exceeded the cap
#include "context.hpp"

namespace insta {
namespace {

constexpr uint32_t kClThreads = 256;
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kClMaxBlocks = 8192;
// a list holds its seed and at most 64 * kSortPerLane more entries (the sort
// kernel's reach); longer ones leave their sub-cell to the BVH
constexpr uint32_t kSortPerLane = 8;
constexpr uint32_t kMaxList = 1u + 64u * kSortPerLane;

__global__ void cl_begin_kernel(uint32_t* __restrict__ state) {
    if (threadIdx.x < kClState) state[threadIdx.x] = 0u;
}

__global__ void cl_invalidate_kernel(uint32_t* __restrict__ state, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) state[kClState * i + 2] = 0u;
}

__global__ void cl_words_kernel(const uint64_t* __restrict__ lin, uint32_t n_words, uint64_t* __restrict__ bits,
                                uint32_t* __restrict__ rank) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_words) return;
    const uint64_t w = lin[i];
    bits[i] = w;
    rank[i] = (uint32_t)__popcll(w);
}

// rank[i] <- popcounts before word i; state[0] = occupied cells,
// state[3] = 1 when they exceed the cap
__global__ __launch_bounds__(kScanThreads) void cl_scan_kernel(uint32_t* __restrict__ rank, uint32_t n_words,
                                                               uint32_t cap, uint32_t* __restrict__ state) {
    __shared__ uint32_t part[kScanThreads];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n_words + kScanThreads - 1) / kScanThreads;
    const uint32_t b = min(t * per, n_words), e = min(b + per, n_words);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += rank[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t c = rank[i];
        rank[i] = run;
        run += c;
    }
    if (t == kScanThreads - 1) {
        state[0] = part[t];
        state[3] = part[t] > cap ? 1u : 0u;
    }
}

__global__ void cl_cells_kernel(const uint64_t* __restrict__ bits, const uint32_t* __restrict__ rank, uint32_t n_words,
                                uint32_t cap, uint32_t* __restrict__ cells) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_words) return;
    uint64_t w = bits[i];
    uint32_t r = rank[i];
    while (w) {
        const uint32_t j = (uint32_t)__ffsll((unsigned long long)w) - 1u;
        if (r < cap) cells[r] = i * 64u + j;
        ++r;
        w &= w - 1ull;
    }
}

struct ClArgs {
    MeshView m;
    const uint32_t* cells;
    uint2* head;
    uint32_t* ent;
    uint32_t* state;
    uint32_t G, S, cap, ent_cap;
};

// sub-cell `sub` of occupied cell `cell`, grown by kCellEps
INSTA_DEV void sub_box(const ClArgs& a, uint32_t cell, uint32_t sub, float4& lo, float4& hi) {
    const uint32_t cx = cell % a.G, cy = (cell / a.G) % a.G, cz = cell / (a.G * a.G);
    const uint32_t sx = sub % a.S, sy = (sub / a.S) % a.S, sz = sub / (a.S * a.S);
    const float inv = 1.0f / (float)(a.G * a.S);
    lo = make_float4((float)(cx * a.S + sx) * inv - kCellEps, (float)(cy * a.S + sy) * inv - kCellEps,
                     (float)(cz * a.S + sz) * inv - kCellEps, 0.0f);
    hi = make_float4((float)(cx * a.S + sx + 1u) * inv + kCellEps, (float)(cy * a.S + sy + 1u) * inv + kCellEps,
                     (float)(cz * a.S + sz + 1u) * inv + kCellEps, 0.0f);
}

INSTA_DEV void tri_verts(const MeshView& m, uint32_t t, V3& a, V3& b, V3& c) {
    const float4 r0 = m.btris[3 * t], r1 = m.btris[3 * t + 1], r2 = m.btris[3 * t + 2];
    a = v3(r0.x, r0.y, r0.z);
    b = v3(r0.w, r1.x, r1.y);
    c = v3(r1.z, r1.w, r2.x);
}

// squared distance between two boxes
INSTA_DEV float box_box_sq(float4 alo, float4 ahi, float4 blo, float4 bhi) {
    const float dx = fmaxf(fmaxf(alo.x - bhi.x, blo.x - ahi.x), 0.0f);
    const float dy = fmaxf(fmaxf(alo.y - bhi.y, blo.y - ahi.y), 0.0f);
    const float dz = fmaxf(fmaxf(alo.z - bhi.z, blo.z - ahi.z), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

// U^2: the largest squared distance from a corner of the box to triangle t,
// with a relative margin over the rounding of the distances and boxes
INSTA_DEV float seed_bound(const MeshView& m, uint32_t t, float4 lo, float4 hi) {
    V3 a, b, c;
    tri_verts(m, t, a, b, c);
    float mx = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const V3 q = v3(k & 1 ? hi.x : lo.x, k & 2 ? hi.y : lo.y, k & 4 ? hi.z : lo.z);
        mx = fmaxf(mx, tri_distance_sq(a, b, c, q));
    }
    return mx * (1.0f + 0x1p-10f) + 0x1p-40f;
}

// Every triangle (BVH order) whose box lies within squared distance U2 of the
// box [lo, hi], in tree order: emit(t, d2).  Depth-first over the implicit
// 4-ary heap; at most 3 pending siblings per inner level.
template <class Emit>
INSTA_DEV void bvh_range(const MeshView& m, float4 lo, float4 hi, float U2, uint32_t* stack, uint32_t stride,
                         Emit&& emit) {
    const uint32_t shift = 2u * m.k;
    int sp = 0;
    uint32_t node = 0;
    for (;;) {
        if (node < m.leaf_base) {
            const uint32_t first = 4u * node + 1u;
            uint32_t next = kNoNode;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint32_t ch = first + c;
                if (box_box_sq(lo, hi, m.nodes[2 * ch], m.nodes[2 * ch + 1]) <= U2) {
                    if (next != kNoNode) stack[(sp++) * stride] = next;
                    next = ch;
                }
            }
            if (next != kNoNode) {
                node = next;
                continue;
            }
        } else {
            const uint32_t j = node - m.leaf_base;
            const uint32_t b = (uint32_t)(((uint64_t)j * m.nf) >> shift);
            const uint32_t e = (uint32_t)(((uint64_t)(j + 1) * m.nf) >> shift);
            for (uint32_t t = b; t < e; ++t) {
                V3 va, vb, vc;
                tri_verts(m, t, va, vb, vc);
                const float4 tlo = make_float4(fminf(fminf(va.x, vb.x), vc.x), fminf(fminf(va.y, vb.y), vc.y),
                                               fminf(fminf(va.z, vb.z), vc.z), 0.0f);
                const float4 thi = make_float4(fmaxf(fmaxf(va.x, vb.x), vc.x), fmaxf(fmaxf(va.y, vb.y), vc.y),
                                               fmaxf(fmaxf(va.z, vb.z), vc.z), 0.0f);
                const float d2 = box_box_sq(lo, hi, tlo, thi);
                if (d2 <= U2) emit(t, d2);
            }
        }
        if (sp == 0) break;
        node = stack[(--sp) * stride];
    }
}

__global__ __launch_bounds__(kClThreads) void cl_count_kernel(ClArgs a) {
    extern __shared__ uint32_t stack[];  // [stack_depth(k)][kClThreads]
    const uint32_t S3 = a.S * a.S * a.S;
    const uint32_t n_sub = min(a.state[0], a.cap) * S3;
    for (uint32_t i = blockIdx.x * kClThreads + threadIdx.x; i < n_sub; i += gridDim.x * kClThreads) {
        const uint32_t cell = a.cells[i / S3];
        float4 lo, hi;
        sub_box(a, cell, i % S3, lo, hi);
        const V3 ctr = v3((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
        const Hit h = bvh_closest(a.m, ctr, stack + threadIdx.x, kClThreads);
        uint2 hd = make_uint2(0u, kNoList);
        if (h.idx >= 0) {
            const float U2 = seed_bound(a.m, (uint32_t)h.idx, lo, hi);
            const uint32_t seed = (uint32_t)h.idx;
            uint32_t count = 1;  // the seed, then every other triangle of the walk
            bvh_range(a.m, lo, hi, U2, stack + threadIdx.x, kClThreads,
                      [&](uint32_t t, float) { count += t != seed ? 1u : 0u; });
            const uint32_t off = atomicAdd(&a.state[1], count);
            if (count <= kMaxList && (uint64_t)off + count <= a.ent_cap)
                hd = make_uint2(off, count | ((uint32_t)h.idx << 16));
        }
        if (hd.y == kNoList) atomicAdd(&a.state[4], 1u);
        a.head[i] = hd;
    }
}

__global__ __launch_bounds__(kClThreads) void cl_fill_kernel(ClArgs a) {
    extern __shared__ uint32_t stack[];  // [stack_depth(k)][kClThreads]
    const uint32_t S3 = a.S * a.S * a.S;
    const uint32_t n_sub = min(a.state[0], a.cap) * S3;
    for (uint32_t i = blockIdx.x * kClThreads + threadIdx.x; i < n_sub; i += gridDim.x * kClThreads) {
        const uint2 hd = a.head[i];
        if (hd.y == kNoList) continue;
        const uint32_t count = hd.y & 0xffffu, seed = hd.y >> 16;
        float4 lo, hi;
        sub_box(a, a.cells[i / S3], i % S3, lo, hi);
        const float U2 = seed_bound(a.m, seed, lo, hi);
        uint32_t* e = a.ent + hd.x;
        e[0] = seed;  // box distance 0: scanned first
        uint32_t k = 1;
        bvh_range(a.m, lo, hi, U2, stack + threadIdx.x, kClThreads, [&](uint32_t t, float d2) {
            if (t != seed && k < count) e[k++] = (__float_as_uint(d2) & 0xffff0000u) | t;
        });
    }
}

// Sort each list's entries after the seed by key (truncated box distance,
// then triangle): one wave per sub-cell, each lane ranks up to kSortPerLane
// keys by counting the smaller ones (keys are distinct), then writes them to
// their ranks.  The keys are read from memory once, into registers (key q of
// lane l is entry l + 64 q); the counting pass broadcasts them from there
// lane by lane (readlane), so it issues no memory reads.  Every store
// follows all of the wave's loads, so no entry is overwritten before it is
// read.
__global__ __launch_bounds__(kClThreads) void cl_sort_kernel(ClArgs a) {
    const uint32_t S3 = a.S * a.S * a.S;
    const uint32_t n_sub = min(a.state[0], a.cap) * S3;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t n_waves = gridDim.x * (kClThreads / 64u);
    for (uint32_t i = blockIdx.x * (kClThreads / 64u) + threadIdx.x / 64u; i < n_sub; i += n_waves) {
        const uint2 hd = a.head[i];
        if (hd.y == kNoList) continue;
        const uint32_t n = (hd.y & 0xffffu) - 1u;
        uint32_t* e = a.ent + hd.x + 1u;
        uint32_t key[kSortPerLane], rank[kSortPerLane];
#pragma unroll
        for (uint32_t q = 0; q < kSortPerLane; ++q) {
            const uint32_t x = lane + 64u * q;
            key[q] = x < n ? e[x] : 0xffffffffu;
            rank[q] = 0u;
        }
        const uint32_t nq = (n + 63u) / 64u;  // key slots in use (wave-uniform)
#pragma unroll
        for (uint32_t qs = 0; qs < kSortPerLane; ++qs) {
            if (qs >= nq) break;
            const uint32_t cnt = min(64u, n - 64u * qs);
            for (uint32_t l = 0; l < cnt; ++l) {
                const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)key[qs], (int)l);
#pragma unroll
                for (uint32_t q = 0; q < kSortPerLane; ++q)
                    if (q < nq) rank[q] += v < key[q] ? 1u : 0u;
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kSortPerLane; ++q)
            if (lane + 64u * q < n) e[rank[q]] = key[q];
    }
}

__global__ void cl_finish_kernel(uint32_t* __restrict__ state) {
    if (threadIdx.x == 0) state[2] = state[3] == 0u ? 1u : 0u;
}

ClArgs cl_args(insta_ctx* ctx, uint32_t slot) {
    ClArgs a;
    const size_t G = ctx->cfg.grid_size, S = ctx->cfg.cell_split;
    a.m = ctx->mesh(slot);
    a.cells = ctx->d_cl_cells + slot * (size_t)ctx->cfg.cell_list_cells;
    a.head = ctx->d_cl_head + slot * (size_t)ctx->cfg.cell_list_cells * S * S * S;
    a.ent = ctx->d_cl_ent + slot * (size_t)ctx->cfg.cell_list_entries;
    a.state = ctx->d_cl_state + kClState * (size_t)slot;
    a.G = (uint32_t)G;
    a.S = (uint32_t)S;
    a.cap = ctx->cfg.cell_list_cells;
    a.ent_cap = ctx->cfg.cell_list_entries;
    return a;
}

}  // namespace

insta_status alloc_cell_lists(insta_ctx* ctx) {
    if (ctx->cfg.cell_split == 0) return INSTA_OK;
    const size_t n = ctx->n_slots, G = ctx->cfg.grid_size, S = ctx->cfg.cell_split;
    const size_t words = G * G * G / 64;
    const size_t cells = ctx->cfg.cell_list_cells, ents = ctx->cfg.cell_list_entries;
    bool ok = true;
    auto alloc = [&](void** p, size_t bytes) {
        if (ok && hipMalloc(p, bytes) != hipSuccess) ok = false;
    };
    alloc((void**)&ctx->d_cl_bits, n * words * sizeof(uint64_t));
    alloc((void**)&ctx->d_cl_rank, n * words * sizeof(uint32_t));
    alloc((void**)&ctx->d_cl_cells, n * cells * sizeof(uint32_t));
    alloc((void**)&ctx->d_cl_head, n * cells * S * S * S * sizeof(uint2));
    alloc((void**)&ctx->d_cl_ent, n * ents * sizeof(uint32_t));
    alloc((void**)&ctx->d_cl_state, n * kClState * sizeof(uint32_t));
    if (!ok) {
        free_cell_lists(ctx);
        set_error("device allocation of the candidate lists failed (lower cell_list_cells / cell_list_entries)");
        return INSTA_ERR_OOM;
    }
    // every slot starts without lists
    if (hipMemset(ctx->d_cl_state, 0, n * kClState * sizeof(uint32_t)) != hipSuccess) {
        free_cell_lists(ctx);
        set_error("candidate list state initialisation failed");
        return INSTA_ERR_HIP;
    }
    return INSTA_OK;
}

void free_cell_lists(insta_ctx* ctx) {
    dev_free(ctx->d_cl_bits);
    dev_free(ctx->d_cl_rank);
    dev_free(ctx->d_cl_cells);
    dev_free(ctx->d_cl_head);
    dev_free(ctx->d_cl_ent);
    dev_free(ctx->d_cl_state);
}

insta_status launch_cell_lists_invalidate(insta_ctx* ctx, uint32_t first, uint32_t n, hipStream_t stream) {
    if (!ctx->d_cl_state || n == 0) return INSTA_OK;
    hipLaunchKernelGGL(cl_invalidate_kernel, dim3((n + 255) / 256), dim3(256), 0, stream,
                       ctx->d_cl_state + kClState * first, n);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_cell_lists(insta_ctx* ctx, uint32_t slot, hipStream_t stream) {
    if (!ctx->d_cl_state) return INSTA_OK;
    const ClArgs a = cl_args(ctx, slot);
    const size_t G = ctx->cfg.grid_size;
    const uint32_t n_words = (uint32_t)(G * G * G / 64);
    uint64_t* bits = ctx->d_cl_bits + slot * (size_t)n_words;
    uint32_t* rank = ctx->d_cl_rank + slot * (size_t)n_words;
    const uint64_t* lin = reinterpret_cast<const uint64_t*>(ctx->grid_lin(slot));
    hipLaunchKernelGGL(cl_begin_kernel, dim3(1), dim3(64), 0, stream, a.state);
    hipLaunchKernelGGL(cl_words_kernel, dim3((n_words + 255) / 256), dim3(256), 0, stream, lin, n_words, bits, rank);
    hipLaunchKernelGGL(cl_scan_kernel, dim3(1), dim3(kScanThreads), 0, stream, rank, n_words, a.cap, a.state);
    hipLaunchKernelGGL(cl_cells_kernel, dim3((n_words + 255) / 256), dim3(256), 0, stream, bits, rank, n_words, a.cap,
                       ctx->d_cl_cells + slot * (size_t)ctx->cfg.cell_list_cells);
    const uint64_t max_sub = (uint64_t)a.cap * a.S * a.S * a.S;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((max_sub + kClThreads - 1) / kClThreads, kClMaxBlocks);
    const size_t shmem = (size_t)stack_depth(ctx->bvh_k) * kClThreads * sizeof(uint32_t);
    hipLaunchKernelGGL(cl_count_kernel, dim3(blocks), dim3(kClThreads), shmem, stream, a);
    hipLaunchKernelGGL(cl_fill_kernel, dim3(blocks), dim3(kClThreads), shmem, stream, a);
    hipLaunchKernelGGL(cl_sort_kernel, dim3(kClMaxBlocks), dim3(kClThreads), 0, stream, a);
    hipLaunchKernelGGL(cl_finish_kernel, dim3(1), dim3(64), 0, stream, a.state);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_mesh_occupancy(insta_ctx* ctx, uint32_t slot, float threshold, hipStream_t stream) {
    insta_status st = launch_occupancy(ctx, slot, threshold, stream);
    if (st != INSTA_OK) return st;
    st = launch_grid_linear(ctx, slot, stream);
    if (st != INSTA_OK) return st;
    return launch_cell_lists(ctx, slot, stream);
}

}  // namespace insta
