// deform.hip -- posed -> canonical sample deformation (the `warp` kernel of
// src/core.cu:442-526 behind Core::surface_closest_point / render_deformed).
//
// Per sample: nearest triangle on the frame's BVH, then either the
// triangle's precomputed inverse frame (most regions) or, on the face region
// (mask 2), an exp-weighted blend of the triangle's and its edge neighbours'
// frames inverted per sample (include/rta/affine.cuh:45-100), followed by the
// canonical triangle's (blended) frame.  Directions go through the 3x3 parts
// and are renormalised.
#include "context.hpp"

namespace insta {
namespace {

constexpr uint32_t kDeformThreads = 256;

struct MeshSet {
    const float4* nodes;
    const float4* btris;
    const float4* proj;
    const float4* proj_inv;
    const float4* cen;
    const int4* adj;
    const uint8_t* built;      // [n_slots]
    size_t node_stride, tri_stride, cen_stride;  // float4 units per slot
    uint32_t nf, leaf_base, k, n_slots, disable_interp;
    uint32_t n_extra;          // extra dims of the density input (0: no selector)
    CellListView cl0;          // candidate lists of slot 0 (null when off)
    size_t cl_words, cl_heads, cl_ents;  // per-slot strides of the list arrays

    INSTA_HD CellListView cells(uint32_t slot) const {
        CellListView c = cl0;
        if (!c.state) return c;
        c.bits += slot * cl_words;
        c.rank += slot * cl_words;
        c.head += slot * cl_heads;
        c.ent += slot * cl_ents;
        c.state += kClState * (size_t)slot;
        return c;
    }

    INSTA_HD MeshView view(uint32_t slot) const {
        MeshView m;
        m.nodes = nodes + slot * node_stride;
        m.btris = btris + slot * tri_stride;
        m.proj = proj + slot * tri_stride;
        m.proj_inv = proj_inv + slot * tri_stride;
        m.cen = cen + slot * cen_stride;
        m.nf = nf;
        m.leaf_base = leaf_base;
        m.k = k;
        return m;
    }
};

MeshSet mesh_set(const insta_ctx* ctx, const uint8_t* built_dev) {
    MeshSet s;
    s.nodes = ctx->d_nodes;
    s.btris = ctx->d_btris;
    s.proj = ctx->d_proj;
    s.proj_inv = ctx->d_proj_inv;
    s.cen = ctx->d_cen;
    s.adj = ctx->d_adj;
    s.built = built_dev;
    s.node_stride = (size_t)ctx->node_cap * 2;
    s.tri_stride = (size_t)ctx->cfg.max_triangles * 3;
    s.cen_stride = ctx->cfg.max_triangles;
    s.nf = ctx->nf;
    s.leaf_base = ctx->leaf_base;
    s.k = ctx->bvh_k;
    s.n_slots = ctx->n_slots;
    s.disable_interp = ctx->cfg.disable_interp;
    s.n_extra = ctx->cfg.n_extra_dims;
    s.cl0 = ctx->cells(0);
    const size_t G = ctx->cfg.grid_size, S = ctx->cfg.cell_split;
    s.cl_words = G * G * G / 64;
    s.cl_heads = (size_t)ctx->cfg.cell_list_cells * S * S * S;
    s.cl_ents = ctx->cfg.cell_list_entries;
    return s;
}

INSTA_DEV void load_rows(const float4* rows, uint32_t id, float m[12]) {
    const float4 r0 = rows[3 * id], r1 = rows[3 * id + 1], r2 = rows[3 * id + 2];
    m[0] = r0.x; m[1] = r0.y; m[2] = r0.z; m[3] = r0.w;
    m[4] = r1.x; m[5] = r1.y; m[6] = r1.z; m[7] = r1.w;
    m[8] = r2.x; m[9] = r2.y; m[10] = r2.z; m[11] = r2.w;
}

// interpolate_transformation without the final inverse
INSTA_DEV void blend_frames(const MeshView& m, const int4 adj, uint32_t id, V3 p, float out[12]) {
    const int32_t nb[3] = {adj.x, adj.y, adj.z};
    float e[3];
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e[k] = 0.0f;
        if (nb[k] >= 0) {
            e[k] = __expf(-4.0f * norm(xyz(m.cen[nb[k]]) - p));
            sum += e[k];
        }
    }
    const float es = __expf(-1.0f * norm(xyz(m.cen[id]) - p));
    sum += es;
#pragma unroll
    for (int i = 0; i < 12; ++i) out[i] = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (nb[k] >= 0) {
            const float w = e[k] / sum;
            float P[12];
            load_rows(m.proj, (uint32_t)nb[k], P);
#pragma unroll
            for (int i = 0; i < 12; ++i) out[i] = out[i] + w * P[i];
        }
    }
    const float w = es / sum;
    float P[12];
    load_rows(m.proj, id, P);
#pragma unroll
    for (int i = 0; i < 12; ++i) out[i] = out[i] + w * P[i];
}

INSTA_DEV void invert_affine(const float m[12], float out[12]) {
    const float a00 = m[0], a01 = m[1], a02 = m[2];
    const float a10 = m[4], a11 = m[5], a12 = m[6];
    const float a20 = m[8], a21 = m[9], a22 = m[10];
    float c00 = a11 * a22 - a12 * a21, c01 = a02 * a21 - a01 * a22, c02 = a01 * a12 - a02 * a11;
    float c10 = a12 * a20 - a10 * a22, c11 = a00 * a22 - a02 * a20, c12 = a02 * a10 - a00 * a12;
    float c20 = a10 * a21 - a11 * a20, c21 = a01 * a20 - a00 * a21, c22 = a00 * a11 - a01 * a10;
    float det = a00 * c00 + a01 * c10 + a02 * c20;
    float id = 1.0f / det;
    float inv[9] = {c00 * id, c01 * id, c02 * id, c10 * id, c11 * id, c12 * id, c20 * id, c21 * id, c22 * id};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        out[r * 4 + 0] = inv[r * 3 + 0];
        out[r * 4 + 1] = inv[r * 3 + 1];
        out[r * 4 + 2] = inv[r * 3 + 2];
        out[r * 4 + 3] = -(inv[r * 3 + 0] * m[3] + inv[r * 3 + 1] * m[7] + inv[r * 3 + 2] * m[11]);
    }
}

INSTA_DEV V3 apply_point(const float m[12], V3 p) {
    return v3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
INSTA_DEV V3 apply_dir(const float m[12], V3 d) {
    return v3(m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
              m[8] * d.x + m[9] * d.y + m[10] * d.z);
}

INSTA_DEV bool in_unit_box(V3 p) {
    return p.x >= 0.0f && p.x <= 1.0f && p.y >= 0.0f && p.y <= 1.0f && p.z >= 0.0f && p.z <= 1.0f;
}

// World-to-triangle (W) and triangle-to-canonical (C) frames of a sample
// whose nearest triangle is `idx` (BVH order on the frame's BVH): the
// triangle's own frames, or on the face region their exp-weighted blend with
// the edge neighbours, the posed blend inverted per sample.
INSTA_DEV void frames_for_hit(const MeshSet& ms, uint32_t frame, int32_t idx, V3 p, float W[12], float C[12],
                              uint32_t& mask) {
    const MeshView fr = ms.view(frame + 1);
    const float4 r2 = fr.btris[3 * idx + 2];
    const uint32_t id = __float_as_uint(r2.y);
    mask = __float_as_uint(r2.z);
    const MeshView cn = ms.view(0);
    if (mask == 2u && !ms.disable_interp) {
        const int4 adj = ms.adj[id];
        float Wb[12];
        blend_frames(fr, adj, id, p, Wb);
        invert_affine(Wb, W);
        blend_frames(cn, adj, id, p, C);
    } else {
        load_rows(fr.proj_inv, id, W);
        load_rows(cn.proj, id, C);
    }
}

// The transform part of `warp` for a sample whose nearest triangle is known.
INSTA_DEV void warp_from_hit(const MeshSet& ms, uint32_t frame, int32_t idx, V3 p, V3 dir_w, V3& out_p, V3& out_d,
                             uint32_t& mask) {
    float W[12], C[12];
    frames_for_hit(ms, frame, idx, p, W, C, mask);
    out_p = apply_point(C, apply_point(W, p));
    const V3 d = v3(dir_w.x * 2.0f - 1.0f, dir_w.y * 2.0f - 1.0f, dir_w.z * 2.0f - 1.0f);
    V3 dc = apply_dir(C, apply_dir(W, d));
    const float n = norm(dc);
    dc = v3(dc.x / n, dc.y / n, dc.z / n);
    out_d = v3((dc.x + 1.0f) * 0.5f, (dc.y + 1.0f) * 0.5f, (dc.z + 1.0f) * 0.5f);
}

// Body of `warp` for one sample; false leaves the sample untouched.  A
// sample in the unit box whose query finds no triangle is counted (the
// reference prints it and returns, src/core.cu:482-485).
INSTA_DEV bool warp_one(const MeshSet& ms, uint32_t frame, V3 p, V3 dir_w, V3& out_p, V3& out_d, uint32_t& mask,
                        uint32_t* stack, uint32_t stride, uint32_t* misses) {
    if (!in_unit_box(p)) return false;
    const Hit h = closest(ms.view(frame + 1), ms.cells(frame + 1), p, stack, stride);
    if (h.idx < 0) {
        atomicAdd(misses, 1u);
        return false;
    }
    warp_from_hit(ms, frame, h.idx, p, dir_w, out_p, out_d, mask);
    return true;
}

// NerfCoordinate samples in place (stride floats per sample)
__global__ __launch_bounds__(kDeformThreads) void deform_aos_kernel(MeshSet ms, float* __restrict__ coords, uint32_t n,
                                                                    uint32_t stride, int32_t frame_arg,
                                                                    uint32_t n_extra, uint32_t* __restrict__ misses) {
    extern __shared__ uint32_t stack[];  // [stack_depth(k)][kDeformThreads]
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* c = coords + (size_t)i * stride;
    const float holder = c[4];
    if (holder < 0.0f) return;
    const uint32_t frame = frame_arg < 0 ? (uint32_t)holder : (uint32_t)frame_arg;
    if (frame >= ms.n_slots - 1 || !ms.built[frame + 1]) return;
    V3 op, od;
    uint32_t mask = 0;
    if (!warp_one(ms, frame, v3(c[0], c[1], c[2]), v3(c[8], c[9], c[10]), op, od, mask, stack + threadIdx.x,
                  kDeformThreads, misses))
        return;
    c[0] = op.x; c[1] = op.y; c[2] = op.z;
    c[8] = od.x; c[9] = od.y; c[10] = od.z;
    if (mask != 7u)
        for (uint32_t m = 0; m < n_extra; ++m) c[stride - n_extra + m] = 1.0f;
}

// Render samples (SoA) in two kernels:
//  hits:  nearest triangle per posed sample (s_pos) into s_hit, one sample
//         per lane, from the sample's sub-cell candidate list when the frame
//         has lists (celllist.hip), else by BVH traversal.  A wave's lanes
//         take the same step of neighbouring rays (an 8x8 pixel tile), so
//         they mostly share a sub-cell and scan the same entries.  Kept
//         apart from the transforms so the traversal runs at low register
//         count, i.e. with many waves resident to cover its chain of
//         dependent node loads.  (Seeding a query with the previous sample's
//         triangle measured slower: the descent reaches a tight bound on its
//         first leaf anyway, and chunking samples per lane breaks the
//         coherence above.)
//  apply: per sample, the frame transforms of `warp` from the hit; writes
//         canonical position (+ mask bits) to s_cpos and the warped
//         direction to s_cdir, whose w selects the extra dims: all 1.0 when
//         the sample was deformed off the lips region (mask 7), else the
//         frame's values (warp only overwrites them in the first case,
//         src/core.cu:514-519).
constexpr uint32_t kMaxBlocks = 65536;

// Frame of a sample: the per-ray table when given (training batches mix
// frames), else the call's frame; -1 leaves the sample undeformed.
INSTA_DEV int32_t sample_frame(int32_t frame, const int32_t* __restrict__ ray_frame, float4 sp) {
    return ray_frame ? ray_frame[__float_as_uint(sp.w)] : frame;
}

// kCount: also sum the traversal work into counts[4] (queries, inner nodes,
// leaves, triangles) -- a separate build for measurement runs
template <bool kCount>
__global__ __launch_bounds__(kDeformThreads) void deform_hits_kernel(MeshSet ms, int32_t frame,
                                                                     const int32_t* __restrict__ ray_frame,
                                                                     const uint32_t* __restrict__ n_dev,
                                                                     const float4* __restrict__ s_pos,
                                                                     int32_t* __restrict__ s_hit,
                                                                     unsigned long long* __restrict__ counts,
                                                                     uint32_t* __restrict__ misses) {
    extern __shared__ uint32_t stack[];  // [stack_depth(k)][kDeformThreads]
    const uint32_t n = *n_dev;
    TraversalCount tc;
    uint32_t queries = 0, missed = 0;
    for (uint32_t i = blockIdx.x * kDeformThreads + threadIdx.x; i < n; i += gridDim.x * kDeformThreads) {
        const float4 sp = s_pos[i];
        const V3 p = v3(sp.x, sp.y, sp.z);
        const int32_t f = sample_frame(frame, ray_frame, sp);
        const bool q = f >= 0 && in_unit_box(p);
        if constexpr (kCount) queries += q ? 1u : 0u;
        const int32_t h = q ? closest_impl<kCount>(ms.view((uint32_t)f + 1), ms.cells((uint32_t)f + 1), p,
                                                   stack + threadIdx.x, kDeformThreads, &tc).idx
                            : -1;
        missed += (q && h < 0) ? 1u : 0u;
        s_hit[i] = h;
    }
    if (misses) {  // one atomic per wave, only when a lane missed
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) missed += __shfl_xor(missed, o);
        if ((threadIdx.x & 63u) == 0 && missed) atomicAdd(misses, missed);
    }
    if constexpr (kCount) {
        unsigned long long v[kTravCounters] = {queries, tc.inner, tc.leaves, tc.triangles, tc.lists, tc.entries};
#pragma unroll
        for (int c = 0; c < kTravCounters; ++c) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
            if ((threadIdx.x & 63u) == 0 && v[c]) atomicAdd(counts + c, v[c]);
        }
    }
}

__global__ __launch_bounds__(kDeformThreads) void deform_apply_kernel(MeshSet ms, int32_t frame,
                                                                      const int32_t* __restrict__ ray_frame,
                                                                      const uint32_t* __restrict__ n_dev,
                                                                      const float4* __restrict__ s_pos,
                                                                      const float4* __restrict__ r_d,
                                                                      const int32_t* __restrict__ s_hit,
                                                                      float4* __restrict__ s_cpos,
                                                                      float4* __restrict__ s_cdir) {
    const uint32_t n = *n_dev;
    for (uint32_t i = blockIdx.x * kDeformThreads + threadIdx.x; i < n; i += gridDim.x * kDeformThreads) {
        const float4 sp = s_pos[i];
        const float4 rd = r_d[__float_as_uint(sp.w)];
        const V3 p = v3(sp.x, sp.y, sp.z);
        const V3 dir_w = v3((rd.x + 1.0f) * 0.5f, (rd.y + 1.0f) * 0.5f, (rd.z + 1.0f) * 0.5f);
        V3 op = p, od = dir_w;
        uint32_t mask = 0;
        const int32_t f = sample_frame(frame, ray_frame, sp);
        const int32_t h = f >= 0 ? s_hit[i] : -1;
        if (h >= 0) warp_from_hit(ms, (uint32_t)f, h, p, dir_w, op, od, mask);
        const bool frame_values = ms.n_extra && f >= 0 && !(h >= 0 && mask != 7u);
        const uint32_t sel = frame_values ? (uint32_t)f + 1u : kExtraOnes;
        s_cpos[i] = make_float4(op.x, op.y, op.z, __uint_as_float(mask));
        s_cdir[i] = make_float4(od.x, od.y, od.z, __uint_as_float(sel));
    }
}

// warp_random_samples (src/core.cu:528-579): positions (stride 3) of
// occupancy-grid samples mapped to canonical space through a random posed
// frame each: the PCG32 stream advanced by i * (n_frames - 1) draws gives
// frame = floor(u * (n_frames - 1)), so the last frame is never chosen (the
// reference's own off-by-one, kept).  No unit-box test, as in the reference.
__global__ __launch_bounds__(kDeformThreads) void deform_positions_kernel(MeshSet ms, float* __restrict__ pos,
                                                                         uint32_t n, uint32_t n_meshes, Pcg32 rng,
                                                                         uint32_t* __restrict__ misses) {
    extern __shared__ uint32_t stack[];  // [stack_depth(k)][kDeformThreads]
    const uint32_t i = blockIdx.x * kDeformThreads + threadIdx.x;
    if (i >= n) return;
    rng.advance((uint64_t)i * n_meshes);
    const uint32_t frame = (uint32_t)(rng.next_float() * (float)n_meshes);
    const V3 p = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    const int32_t h = closest(ms.view(frame + 1), ms.cells(frame + 1), p, stack + threadIdx.x, kDeformThreads).idx;
    if (h < 0) {  // src/core.cu:554-557
        atomicAdd(misses, 1u);
        return;
    }
    float W[12], C[12];
    uint32_t mask;
    frames_for_hit(ms, frame, h, p, W, C, mask);
    const V3 c = apply_point(C, apply_point(W, p));
    pos[3 * i] = c.x;
    pos[3 * i + 1] = c.y;
    pos[3 * i + 2] = c.z;
}

uint32_t blocks_for(uint64_t items) {
    const uint64_t b = (items + kDeformThreads - 1) / kDeformThreads;
    return (uint32_t)(b < kMaxBlocks ? (b ? b : 1) : kMaxBlocks);
}

}  // namespace

insta_status launch_deform_aos(insta_ctx* ctx, float* coords, uint32_t n, uint32_t stride, int32_t frame,
                               hipStream_t stream) {
    if (n == 0) return INSTA_OK;
    const size_t shmem = (size_t)stack_depth(ctx->bvh_k) * kDeformThreads * sizeof(uint32_t);
    hipLaunchKernelGGL(deform_aos_kernel, dim3((n + kDeformThreads - 1) / kDeformThreads), dim3(kDeformThreads), shmem,
                       stream, mesh_set(ctx, ctx->d_built), coords, n, stride, frame, ctx->cfg.n_extra_dims,
                       ctx->d_misses);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_deform_samples(insta_ctx* ctx, int32_t frame, const int32_t* ray_frame, const uint32_t* n_dev,
                                   uint32_t n_max, const float4* s_pos, const float4* r_d, int32_t* s_hit,
                                   float4* s_cpos, float4* s_cdir, uint32_t* miss_counter, hipStream_t stream) {
    if (n_max == 0) return INSTA_OK;
    const MeshSet ms = mesh_set(ctx, ctx->d_built);
    if (frame >= 0 || ray_frame) {
        const size_t shmem = (size_t)stack_depth(ctx->bvh_k) * kDeformThreads * sizeof(uint32_t);
        if (ctx->d_trav_counts)
            hipLaunchKernelGGL(deform_hits_kernel<true>, dim3(blocks_for(n_max)), dim3(kDeformThreads), shmem, stream,
                               ms, frame, ray_frame, n_dev, s_pos, s_hit, ctx->d_trav_counts, miss_counter);
        else
            hipLaunchKernelGGL(deform_hits_kernel<false>, dim3(blocks_for(n_max)), dim3(kDeformThreads), shmem,
                               stream, ms, frame, ray_frame, n_dev, s_pos, s_hit, nullptr, miss_counter);
        INSTA_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(deform_apply_kernel, dim3(blocks_for(n_max)), dim3(kDeformThreads), 0, stream, ms, frame,
                       ray_frame, n_dev, s_pos, r_d, s_hit, s_cpos, s_cdir);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_deform_positions(insta_ctx* ctx, float* pos, uint32_t n, uint32_t n_frames, const Pcg32& rng,
                                     hipStream_t stream) {
    if (n == 0) return INSTA_OK;
    const size_t shmem = (size_t)stack_depth(ctx->bvh_k) * kDeformThreads * sizeof(uint32_t);
    hipLaunchKernelGGL(deform_positions_kernel, dim3((n + kDeformThreads - 1) / kDeformThreads), dim3(kDeformThreads),
                       shmem, stream, mesh_set(ctx, ctx->d_built), pos, n, n_frames - 1, rng, ctx->d_misses);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

insta_status launch_deform_soa(insta_ctx* ctx, RenderWS& ws, int32_t frame, const uint32_t* n_dev, uint32_t n_max,
                               uint32_t* miss_counter, hipStream_t stream) {
    return launch_deform_samples(ctx, frame, nullptr, n_dev, n_max, ws.s_pos, ws.r_d, ws.s_hit, ws.s_cpos, ws.s_cdir,
                                 miss_counter, stream);
}

}  // namespace insta
