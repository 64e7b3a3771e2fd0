// capi.hip -- the extern "C" boundary declared in include/insta_capi.h.
//
// Host-side orchestration only: validation, allocation at creation time,
// adjacency/render-box set-up from the canonical mesh, and dispatch to the
// kernel launchers.  No exception crosses the ABI; failures return a status
// and leave a message for insta_last_error().
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "context.hpp"

namespace insta {


namespace {
thread_local std::string g_error;

__global__ void set_flags_kernel(uint8_t* flags, uint32_t first, uint32_t n, uint8_t value) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[first + i] = value;
}

__global__ void export_order_kernel(const float4* __restrict__ btris, uint32_t nf, uint32_t* __restrict__ order) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nf) order[i] = __float_as_uint(btris[3 * i + 2].y);
}

// inner levels of the 4-ary tree: the largest k with 4^k * leaf_min <= n, so
// the 4^k count-balanced leaves hold leaf_min .. 4 * leaf_min - 1 triangles
// (k = 0 below leaf_min)
uint32_t bvh_levels(uint32_t n, uint32_t leaf_min) {
    uint32_t k = 0;
    while (((uint64_t)leaf_min << (2 * (k + 1))) <= n) ++k;
    return k;
}

// Hash-grid and MLP parameter layout.  The per-level scale is evaluated in
// double and rounded to float once, so the kernels and any CPU check see the
// same float scale.
void compute_layout(const insta_config& c, insta_layout& L) {
    std::memset(&L, 0, sizeof(L));
    const uint64_t T = uint64_t(1) << c.log2_hashmap_size;
    uint64_t off = 0;
    for (uint32_t l = 0; l < c.n_levels; ++l) {
        const double s = std::exp2(l * std::log2((double)c.per_level_scale)) * c.base_resolution - 1.0;
        const float sf = (float)s;
        const uint32_t res = (uint32_t)std::ceil(sf) + 1u;
        const uint64_t verts = uint64_t(res) * res * res;
        const bool dense = verts <= T;
        const uint64_t size = dense ? ((verts + 7) / 8) * 8 : T;
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
    const uint64_t w = c.n_neurons;
    uint64_t n_mlp = w * L.density_in + (uint64_t)(c.density_hidden_layers - 1) * w * w + c.density_out * w;
    n_mlp += w * L.rgb_in + (uint64_t)(c.rgb_hidden_layers - 1) * w * w + 3 * w;
    L.n_mlp_params = n_mlp;
    L.n_params = n_mlp + L.n_grid_params;
}

bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }

insta_status validate(const insta_config& c) {
    auto bad = [](const char* m) {
        set_error(m);
        return INSTA_ERR_INVALID;
    };
    if (c.n_levels == 0 || c.n_levels > INSTA_MAX_LEVELS) return bad("n_levels must be in [1, 32]");
    if (!(c.n_features_per_level == 1 || c.n_features_per_level == 2 || c.n_features_per_level == 4 ||
          c.n_features_per_level == 8))
        return bad("n_features_per_level must be 1, 2, 4 or 8");
    if (c.log2_hashmap_size < 4 || c.log2_hashmap_size > 26) return bad("log2_hashmap_size must be in [4, 26]");
    if (c.base_resolution < 2 || !(c.per_level_scale >= 1.0f)) return bad("base_resolution >= 2, per_level_scale >= 1");
    if (c.n_neurons == 0 || c.n_neurons > 256 || c.density_out == 0 || c.density_out > 48)
        return bad("n_neurons in [1, 256], density_out in [1, 48]");
    if (c.density_hidden_layers == 0 || c.rgb_hidden_layers == 0) return bad("at least one hidden layer per MLP");
    if (c.sh_degree != 4) return bad("sh_degree must be 4");
    if (c.n_extra_dims != 0 && c.n_extra_dims != 16) return bad("n_extra_dims must be 0 or 16");
    if (!is_pow2(c.grid_size) || c.grid_size < 8 || c.grid_size > 1024) return bad("grid_size: power of two in [8, 1024]");
    if (!(c.dt > 0.0f) || c.max_steps == 0) return bad("dt > 0 and max_steps >= 1");
    if (!(c.min_transmittance >= 0.0f && c.min_transmittance < 1.0f)) return bad("min_transmittance in [0, 1)");
    if (c.max_frames == 0 || c.max_frames > 4096) return bad("max_frames in [1, 4096]");
    if (c.max_triangles == 0 || c.max_triangles > INSTA_MAX_TRIANGLES) return bad("max_triangles in [1, 16384]");
    if (c.max_vertices == 0 || c.max_vertices > (1u << 24)) return bad("max_vertices in [1, 2^24]");
    if (c.bvh_leaf_min == 0 || c.bvh_leaf_min > 64) return bad("bvh_leaf_min in [1, 64]");
    if (c.precision > INSTA_PRECISION_FP32_SCALAR) return bad("unknown precision");
    if (!(c.cell_split == 0 || c.cell_split == 1 || c.cell_split == 2 || c.cell_split == 4))
        return bad("cell_split must be 0, 1, 2 or 4");
    if (c.cell_split && (c.cell_list_cells == 0 || c.cell_list_cells > c.grid_size * c.grid_size * c.grid_size ||
                         c.cell_list_entries == 0 || c.cell_list_entries > (1u << 28)))
        return bad("cell_list_cells in [1, grid_size^3], cell_list_entries in [1, 2^28]");
    return INSTA_OK;
}

#define INSTA_CHECK_CTX(ctx)                       \
    if (!(ctx)) {                                  \
        insta::set_error("null context");          \
        return INSTA_ERR_INVALID;                  \
    }

insta_status mark_built(insta_ctx* ctx, uint32_t first, uint32_t n, uint8_t v, hipStream_t stream) {
    hipLaunchKernelGGL(set_flags_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, ctx->d_built, first, n, v);
    INSTA_HIP(hipGetLastError());
    for (uint32_t i = 0; i < n; ++i) ctx->slot_built[first + i] = v;
    return INSTA_OK;
}

}  // namespace

// device side of mark_built only (capturable into a graph)
insta_status launch_set_built(insta_ctx* ctx, uint32_t first, uint32_t n, hipStream_t stream) {
    hipLaunchKernelGGL(set_flags_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, ctx->d_built, first, n,
                       (uint8_t)1);
    INSTA_HIP(hipGetLastError());
    return INSTA_OK;
}

namespace {

bool slot_of(const insta_ctx* ctx, int32_t frame, uint32_t& slot) {
    if (frame < -1 || frame >= (int32_t)ctx->cfg.max_frames) return false;
    slot = (uint32_t)(frame + 1);
    return ctx->slot_built[slot] != 0;
}

}  // namespace

insta_status check_render_args(const insta_ctx* ctx, const insta_camera* cam, uint32_t width, uint32_t height,
                               int32_t frame, int32_t occupancy, const float* rgba) {
    if (!cam || !rgba || width == 0 || height == 0 || (uint64_t)width * height > (1ull << 26)) {
        set_error("render: bad camera, output pointer or resolution");
        return INSTA_ERR_INVALID;
    }
    if (!ctx->params_set) {
        set_error("render: parameters not set");
        return INSTA_ERR_STATE;
    }
    if (!mfma_network_supported(ctx->cfg) && (!network_supported(ctx->cfg) || !scalar_precision(ctx->cfg))) {
        set_error("render: configuration not supported by the fused MLP kernels");
        return INSTA_ERR_UNSUPPORTED;
    }
    uint32_t slot;
    if (!slot_of(ctx, frame, slot)) {
        set_error("render: frame not built");
        return INSTA_ERR_STATE;
    }
    if (occupancy < -1 || occupancy >= (int32_t)ctx->cfg.max_frames || !ctx->grid_built[occupancy + 1]) {
        set_error("render: occupancy slot not built");
        return INSTA_ERR_STATE;
    }
    return INSTA_OK;
}

void set_error(const std::string& msg) { g_error = msg; }

insta_status hip_fail(hipError_t e, const char* what) {
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? INSTA_ERR_OOM : INSTA_ERR_HIP;
}

}  // namespace insta

using insta::set_error;

extern "C" {

const char* insta_last_error(void) { return insta::g_error.c_str(); }

const char* insta_version(void) { return "insta-amd 0.1 (gfx950)"; }

void insta_default_config(insta_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->n_levels = 16;
    c->n_features_per_level = 2;
    c->log2_hashmap_size = 19;
    c->base_resolution = 16;
    c->per_level_scale = (float)std::exp(std::log(2048.0 / 16.0) / 15.0);
    c->n_neurons = 64;
    c->density_hidden_layers = 1;
    c->density_out = 16;
    c->rgb_hidden_layers = 2;
    c->sh_degree = 4;
    c->n_extra_dims = 0;
    c->density_logit_bias = 1.0f;
    c->grid_size = 128;
    c->dt = 1.73205080757f / 1024.0f;
    c->max_steps = 1024;
    c->min_transmittance = 0.01f;
    c->disable_interp = 0;
    c->precision = INSTA_PRECISION_FP32;
    c->max_frames = 16;
    c->max_triangles = INSTA_MAX_TRIANGLES;
    c->max_vertices = 1u << 16;
    c->bvh_leaf_min = 4;
    c->cell_split = 4;
    c->cell_list_cells = 1u << 16;
    c->cell_list_entries = 1u << 25;
}

insta_status insta_query_layout(const insta_config* cfg, insta_layout* out) {
    if (!cfg || !out) {
        set_error("null argument");
        return INSTA_ERR_INVALID;
    }
    insta_status st = insta::validate(*cfg);
    if (st != INSTA_OK) return st;
    insta::compute_layout(*cfg, *out);
    return INSTA_OK;
}

insta_status insta_create(const insta_config* cfg, insta_ctx** out) {
    if (!cfg || !out) {
        set_error("null argument");
        return INSTA_ERR_INVALID;
    }
    *out = nullptr;
    insta_status st = insta::validate(*cfg);
    if (st != INSTA_OK) return st;
    insta_ctx* ctx = new (std::nothrow) insta_ctx();
    if (!ctx) {
        set_error("host allocation failed");
        return INSTA_ERR_OOM;
    }
    ctx->cfg = *cfg;
    insta::compute_layout(ctx->cfg, ctx->layout);
    if (hipGetDevice(&ctx->device) != hipSuccess) {
        delete ctx;
        set_error("no HIP device");
        return INSTA_ERR_HIP;
    }
    const uint32_t kmax = insta::bvh_levels(cfg->max_triangles, 1);  // the deepest tree any leaf size needs
    ctx->node_cap = ((1u << (2 * (kmax + 1))) - 1) / 3;
    ctx->n_slots = cfg->max_frames + 1;
    ctx->slot_built.assign(ctx->n_slots, 0);
    ctx->grid_built.assign(ctx->n_slots, 0);
    const size_t S = ctx->n_slots, T = cfg->max_triangles, G = cfg->grid_size;
    ctx->scratch_bytes = std::max<size_t>((size_t)cfg->max_vertices * 3 * sizeof(float), 1 << 16) + 4096;
    bool ok = true;
    auto alloc = [&](void** p, size_t bytes) {
        if (ok && hipMalloc(p, bytes) != hipSuccess) ok = false;
    };
    alloc((void**)&ctx->d_faces, T * 3 * sizeof(int32_t));
    alloc((void**)&ctx->d_masks, T * sizeof(uint32_t));
    alloc((void**)&ctx->d_adj, T * sizeof(int4));
    alloc((void**)&ctx->d_verts, S * cfg->max_vertices * 3 * sizeof(float));
    alloc((void**)&ctx->d_nodes, S * ctx->node_cap * 2 * sizeof(float4));
    alloc((void**)&ctx->d_btris, S * T * 3 * sizeof(float4));
    alloc((void**)&ctx->d_proj, S * T * 3 * sizeof(float4));
    alloc((void**)&ctx->d_proj_inv, S * T * 3 * sizeof(float4));
    alloc((void**)&ctx->d_cen, S * T * sizeof(float4));
    alloc((void**)&ctx->d_grids, S * (G * G * G / 8));
    alloc((void**)&ctx->d_grids_lin, S * (G * G * G / 8));
    alloc((void**)&ctx->d_built, S);
    alloc((void**)&ctx->d_build_lists, S * 6 * T * sizeof(uint16_t));
    alloc((void**)&ctx->d_build_q, S * 3 * T * sizeof(uint32_t));
    alloc((void**)&ctx->d_params, ctx->layout.n_params * sizeof(float));
    alloc(&ctx->d_wfrag, 64 * 1024 * sizeof(float));
    if (cfg->precision == INSTA_PRECISION_FP16 || cfg->precision == INSTA_PRECISION_BF16)
        alloc(&ctx->d_table_h, ctx->layout.n_grid_params * 2);
    alloc(&ctx->d_scratch, ctx->scratch_bytes);
    alloc((void**)&ctx->d_misses, sizeof(uint32_t));
    if (ok && hipMemset(ctx->d_misses, 0, sizeof(uint32_t)) != hipSuccess) ok = false;
    if (cfg->n_extra_dims) {
        alloc((void**)&ctx->d_frame_extra, (size_t)cfg->max_frames * cfg->n_extra_dims * sizeof(float));
        alloc((void**)&ctx->d_extra_bias, S * 64 * sizeof(float));
        if (ok && hipMemset(ctx->d_frame_extra, 0, (size_t)cfg->max_frames * cfg->n_extra_dims * sizeof(float)) !=
                      hipSuccess)
            ok = false;
        if (ok && hipMemset(ctx->d_extra_bias, 0, S * 64 * sizeof(float)) != hipSuccess) ok = false;
    }
    if (ok && hipMemset(ctx->d_built, 0, S) != hipSuccess) ok = false;
    if (ok && hipMemset(ctx->d_grids, 0, S * (G * G * G / 8)) != hipSuccess) ok = false;
    if (ok && hipMemset(ctx->d_grids_lin, 0, S * (G * G * G / 8)) != hipSuccess) ok = false;
    if (ok && insta::noalloc_prepare_device(ctx->device) != INSTA_OK) {
        const std::string msg = insta_last_error();
        insta_destroy(ctx);
        set_error(msg);
        return INSTA_ERR_HIP;
    }
    if (ok && insta::alloc_cell_lists(ctx) != INSTA_OK) {
        const std::string msg = insta_last_error();
        insta_destroy(ctx);
        set_error(msg);
        return INSTA_ERR_OOM;
    }
    if (!ok) {
        insta_destroy(ctx);
        set_error("device allocation failed in insta_create");
        return INSTA_ERR_OOM;
    }
    *out = ctx;
    return INSTA_OK;
}

void insta_destroy(insta_ctx* ctx) {
    if (!ctx) return;
    insta::free_workspace(ctx);
    insta::free_train(ctx);
    insta::free_density_grid(ctx);
    using insta::dev_free;
    dev_free(ctx->d_faces); dev_free(ctx->d_masks); dev_free(ctx->d_adj); dev_free(ctx->d_verts);
    dev_free(ctx->d_nodes); dev_free(ctx->d_btris); dev_free(ctx->d_proj); dev_free(ctx->d_proj_inv);
    dev_free(ctx->d_cen); dev_free(ctx->d_grids); dev_free(ctx->d_grids_lin); dev_free(ctx->d_built); dev_free(ctx->d_params);
    dev_free(ctx->d_build_lists); dev_free(ctx->d_build_q);
    dev_free(ctx->d_scratch); dev_free(ctx->d_wfrag); dev_free(ctx->d_table_h);
    dev_free(ctx->d_frame_extra); dev_free(ctx->d_extra_bias); dev_free(ctx->d_misses);
    insta::free_cell_lists(ctx);
    delete ctx;
}

insta_status insta_set_canonical_mesh(insta_ctx* ctx, const float* verts, uint32_t nv, const int32_t* faces,
                                      uint32_t nf, const uint32_t* mask_ids, void* stream_) {
    INSTA_CHECK_CTX(ctx);
    hipStream_t stream = (hipStream_t)stream_;
    if (!verts || !faces || !mask_ids || nv == 0 || nf == 0) {
        set_error("set_canonical_mesh: empty mesh");
        return INSTA_ERR_INVALID;
    }
    if (nv > ctx->cfg.max_vertices || nf > ctx->cfg.max_triangles) {
        set_error("set_canonical_mesh: mesh exceeds max_vertices / max_triangles");
        return INSTA_ERR_INVALID;
    }
    for (uint64_t i = 0; i < (uint64_t)nf * 3; ++i) {
        if (faces[i] < 0 || (uint32_t)faces[i] >= nv) {
            set_error("set_canonical_mesh: face index out of range");
            return INSTA_ERR_INVALID;
        }
    }
    // edge adjacency (src/tiny_mesh.cu:316-371): faces sharing an edge,
    // ascending face id, self removed, -1 padded
    std::vector<std::vector<int32_t>> vf(nv);
    for (uint32_t i = 0; i < nf; ++i)
        for (int k = 0; k < 3; ++k) vf[faces[3 * i + k]].push_back((int32_t)i);
    std::vector<int4> adj(nf);
    for (uint32_t i = 0; i < nf; ++i) {
        std::vector<int32_t> ids;
        const int32_t v[3] = {faces[3 * i], faces[3 * i + 1], faces[3 * i + 2]};
        const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (const auto& pr : pairs) {
            const auto& A = vf[v[pr[0]]];
            const auto& B = vf[v[pr[1]]];
            for (int32_t f : A)
                if (std::find(B.begin(), B.end(), f) != B.end()) ids.push_back(f);
        }
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        int32_t nb[3] = {-1, -1, -1};
        int k = 0;
        for (int32_t f : ids) {
            if (f == (int32_t)i) continue;
            if (k < 3) nb[k] = f;
            ++k;
        }
        adj[i] = make_int4(nb[0], nb[1], nb[2], -1);
    }
    // render box: vertex AABB inflated by 0.15 |diag| (tiny_mesh.cu:157-162),
    // clipped to the unit cube (testbed_nerf.cu:2675-2682)
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::fmin(lo[d], verts[3 * i + d]);
            hi[d] = std::fmax(hi[d], verts[3 * i + d]);
        }
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    const float amount = std::sqrt(dx * dx + dy * dy + dz * dz) * 0.15f;
    for (int d = 0; d < 3; ++d) {
        ctx->render_lo[d] = std::fmax(lo[d] - amount, 0.0f);
        ctx->render_hi[d] = std::fmin(hi[d] + amount, 1.0f);
    }
    ctx->nv = nv;
    ctx->nf = nf;
    ctx->bvh_k = insta::bvh_levels(nf, ctx->cfg.bvh_leaf_min);
    ctx->leaf_base = ((1u << (2 * ctx->bvh_k)) - 1) / 3;
    ctx->n_nodes = ctx->leaf_base + (1u << (2 * ctx->bvh_k));
    insta_status st = insta::mark_built(ctx, 0, ctx->n_slots, 0, stream);
    if (st != INSTA_OK) return st;
    std::fill(ctx->grid_built.begin(), ctx->grid_built.end(), 0);
    INSTA_HIP(hipMemcpyAsync(ctx->d_faces, faces, (size_t)nf * 3 * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    INSTA_HIP(hipMemcpyAsync(ctx->d_masks, mask_ids, (size_t)nf * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    INSTA_HIP(hipMemcpyAsync(ctx->d_adj, adj.data(), (size_t)nf * sizeof(int4), hipMemcpyHostToDevice, stream));
    float* tmp = static_cast<float*>(ctx->d_scratch);
    INSTA_HIP(hipMemcpyAsync(tmp, verts, (size_t)nv * 3 * sizeof(float), hipMemcpyHostToDevice, stream));
    st = insta::launch_build_slots(ctx, 0, 1, tmp, stream);
    if (st != INSTA_OK) return st;
    st = insta::mark_built(ctx, 0, 1, 1, stream);
    if (st != INSTA_OK) return st;
    // host arrays above were staged by the copies; wait so they may be freed
    INSTA_HIP(hipStreamSynchronize(stream));
    return INSTA_OK;
}

insta_status insta_set_frames(insta_ctx* ctx, uint32_t first_frame, uint32_t n_frames, const float* verts_dev,
                              void* stream_) {
    INSTA_CHECK_CTX(ctx);
    hipStream_t stream = (hipStream_t)stream_;
    if (!ctx->slot_built[0]) {
        set_error("set_frames: canonical mesh not set");
        return INSTA_ERR_STATE;
    }
    if (!verts_dev || n_frames == 0 || (uint64_t)first_frame + n_frames > ctx->cfg.max_frames) {
        set_error("set_frames: frame range outside [0, max_frames)");
        return INSTA_ERR_INVALID;
    }
    insta_status st = insta::launch_build_slots(ctx, first_frame + 1, n_frames, verts_dev, stream);
    if (st != INSTA_OK) return st;
    for (uint32_t f = 0; f < n_frames; ++f) ctx->grid_built[first_frame + 1 + f] = 0;
    return insta::mark_built(ctx, first_frame + 1, n_frames, 1, stream);
}

insta_status insta_set_frame_expression(insta_ctx* ctx, uint32_t frame, const float* bases_dev, uint32_t n_bases,
                                        const float* coeffs_dev, void* stream_) {
    INSTA_CHECK_CTX(ctx);
    hipStream_t stream = (hipStream_t)stream_;
    if (!ctx->slot_built[0]) {
        set_error("set_frame_expression: canonical mesh not set");
        return INSTA_ERR_STATE;
    }
    if (frame >= ctx->cfg.max_frames || !bases_dev || !coeffs_dev) {
        set_error("set_frame_expression: bad frame or null pointer");
        return INSTA_ERR_INVALID;
    }
    insta_status st = insta::launch_blend_expression(ctx, frame + 1, bases_dev, n_bases, coeffs_dev, stream);
    if (st != INSTA_OK) return st;
    st = insta::launch_extra_from_coeffs(ctx, frame, coeffs_dev, n_bases, stream);
    if (st != INSTA_OK) return st;
    ctx->grid_built[frame + 1] = 0;
    return insta::mark_built(ctx, frame + 1, 1, 1, stream);
}

insta_status insta_set_frame_extra_dims(insta_ctx* ctx, uint32_t first_frame, uint32_t n_frames,
                                        const float* values_dev, void* stream_) {
    INSTA_CHECK_CTX(ctx);
    hipStream_t stream = (hipStream_t)stream_;
    const uint32_t X = ctx->cfg.n_extra_dims;
    if (X == 0) {
        set_error("set_frame_extra_dims: n_extra_dims is 0");
        return INSTA_ERR_STATE;
    }
    if (n_frames && !values_dev) {
        set_error("set_frame_extra_dims: null pointer");
        return INSTA_ERR_INVALID;
    }
    if ((uint64_t)first_frame + n_frames > ctx->cfg.max_frames) {
        set_error("set_frame_extra_dims: frame range outside [0, max_frames)");
        return INSTA_ERR_INVALID;
    }
    if (n_frames == 0) return INSTA_OK;
    INSTA_HIP(hipMemcpyAsync(ctx->d_frame_extra + (size_t)first_frame * X, values_dev,
                             (size_t)n_frames * X * sizeof(float), hipMemcpyDeviceToDevice, stream));
    return insta::update_extra_rows(ctx, first_frame + 1, n_frames, stream);
}

insta_status insta_closest_triangles(insta_ctx* ctx, int32_t frame, const float* points_dev, uint32_t n,
                                     int32_t* ids_dev, float* dist_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    uint32_t slot;
    if (!insta::slot_of(ctx, frame, slot)) {
        set_error("closest_triangles: frame not built");
        return INSTA_ERR_STATE;
    }
    if (n && (!points_dev || !ids_dev)) {
        set_error("closest_triangles: null pointer");
        return INSTA_ERR_INVALID;
    }
    return insta::launch_closest(ctx, slot, points_dev, n, ids_dev, dist_dev, true, (hipStream_t)stream);
}

insta_status insta_get_cell_list_info(insta_ctx* ctx, int32_t frame, insta_cell_list_info* out, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!out || frame < -1 || frame >= (int32_t)ctx->cfg.max_frames) {
        set_error("get_cell_list_info: bad frame or null pointer");
        return INSTA_ERR_INVALID;
    }
    *out = insta_cell_list_info{};
    if (!ctx->d_cl_state) return INSTA_OK;
    uint32_t st[insta::kClState];
    const uint32_t slot = (uint32_t)(frame + 1);
    INSTA_HIP(hipMemcpyAsync(st, ctx->d_cl_state + insta::kClState * (size_t)slot, sizeof(st), hipMemcpyDeviceToHost,
                             (hipStream_t)stream));
    INSTA_HIP(hipStreamSynchronize((hipStream_t)stream));
    out->occupied_cells = st[0];
    out->entries = st[1];
    out->valid = st[2];
    out->subcells_without_list = st[4];
    return INSTA_OK;
}

insta_status insta_export_bvh(insta_ctx* ctx, int32_t frame, float* nodes_dev, uint32_t* order_dev,
                              uint32_t* n_nodes, void* stream_) {
    INSTA_CHECK_CTX(ctx);
    hipStream_t stream = (hipStream_t)stream_;
    uint32_t slot;
    if (!insta::slot_of(ctx, frame, slot)) {
        set_error("export_bvh: frame not built");
        return INSTA_ERR_STATE;
    }
    if (n_nodes) *n_nodes = ctx->n_nodes;
    const insta::MeshView m = ctx->mesh(slot);
    if (nodes_dev)
        INSTA_HIP(hipMemcpyAsync(nodes_dev, m.nodes, (size_t)ctx->n_nodes * 2 * sizeof(float4),
                                 hipMemcpyDeviceToDevice, stream));
    if (order_dev) {
        hipLaunchKernelGGL(insta::export_order_kernel, dim3((ctx->nf + 255) / 256), dim3(256), 0, stream, m.btris,
                           ctx->nf, order_dev);
        INSTA_HIP(hipGetLastError());
    }
    return INSTA_OK;
}

insta_status insta_deform_samples(insta_ctx* ctx, float* coords_dev, uint32_t n, uint32_t stride_floats,
                                  int32_t frame, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!ctx->slot_built[0]) {
        set_error("deform_samples: canonical mesh not set");
        return INSTA_ERR_STATE;
    }
    if (stride_floats < 11 + ctx->cfg.n_extra_dims || (n && !coords_dev)) {
        set_error("deform_samples: stride must be >= 11 + n_extra_dims");
        return INSTA_ERR_INVALID;
    }
    if (frame >= (int32_t)ctx->cfg.max_frames || frame < -1) {
        set_error("deform_samples: frame out of range");
        return INSTA_ERR_INVALID;
    }
    return insta::launch_deform_aos(ctx, coords_dev, n, stride_floats, frame, (hipStream_t)stream);
}

insta_status insta_build_occupancy(insta_ctx* ctx, int32_t frame, float threshold, uint8_t* bitfield_dev,
                                   void* stream_) {
    INSTA_CHECK_CTX(ctx);
    hipStream_t stream = (hipStream_t)stream_;
    uint32_t slot;
    if (!insta::slot_of(ctx, frame, slot)) {
        set_error("build_occupancy: frame not built");
        return INSTA_ERR_STATE;
    }
    insta_status st = insta::launch_mesh_occupancy(ctx, slot, threshold, stream);
    if (st != INSTA_OK) return st;
    ctx->grid_built[slot] = 1;
    const size_t G = ctx->cfg.grid_size;
    if (bitfield_dev)
        INSTA_HIP(hipMemcpyAsync(bitfield_dev, ctx->grid(slot), G * G * G / 8, hipMemcpyDeviceToDevice, stream));
    return INSTA_OK;
}

insta_status insta_set_occupancy(insta_ctx* ctx, int32_t slot_arg, const uint8_t* bitfield_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (slot_arg < -1 || slot_arg >= (int32_t)ctx->cfg.max_frames || !bitfield_dev) {
        set_error("set_occupancy: bad slot or null pointer");
        return INSTA_ERR_INVALID;
    }
    const uint32_t slot = (uint32_t)(slot_arg + 1);
    const size_t G = ctx->cfg.grid_size;
    INSTA_HIP(hipMemcpyAsync(ctx->grid(slot), bitfield_dev, G * G * G / 8, hipMemcpyDeviceToDevice,
                             (hipStream_t)stream));
    const insta_status st = insta::launch_grid_linear(ctx, slot, (hipStream_t)stream);
    if (st != INSTA_OK) return st;
    ctx->grid_built[slot] = 1;
    return INSTA_OK;
}

insta_status insta_set_params(insta_ctx* ctx, const float* params_dev, uint64_t n_params, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!params_dev || n_params != ctx->layout.n_params) {
        set_error("set_params: n_params does not match the layout");
        return INSTA_ERR_INVALID;
    }
    INSTA_HIP(hipMemcpyAsync(ctx->d_params, params_dev, n_params * sizeof(float), hipMemcpyDeviceToDevice,
                             (hipStream_t)stream));
    insta::reset_inference_average(ctx);
    insta_status st = insta::prepare_network(ctx, (hipStream_t)stream);
    if (st != INSTA_OK) return st;
    ctx->params_set = true;
    return INSTA_OK;
}

insta_status insta_encode(insta_ctx* ctx, const float* positions_dev, uint32_t n, float* features_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!ctx->params_set) {
        set_error("encode: parameters not set");
        return INSTA_ERR_STATE;
    }
    if (n && (!positions_dev || !features_dev)) {
        set_error("encode: null pointer");
        return INSTA_ERR_INVALID;
    }
    return insta::launch_encode(ctx, positions_dev, n, features_dev, (hipStream_t)stream);
}

insta_status insta_encode_backward(insta_ctx* ctx, const float* positions_dev, uint32_t n,
                                   const float* dL_dfeatures_dev, float* grad_tables_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (n && (!positions_dev || !dL_dfeatures_dev || !grad_tables_dev)) {
        set_error("encode_backward: null pointer");
        return INSTA_ERR_INVALID;
    }
    return insta::launch_encode_backward(ctx, positions_dev, n, dL_dfeatures_dev, grad_tables_dev,
                                         (hipStream_t)stream);
}

insta_status insta_network_inference(insta_ctx* ctx, const float* coords_dev, uint32_t n, uint32_t stride_floats,
                                     float* out_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!ctx->params_set) {
        set_error("network_inference: parameters not set");
        return INSTA_ERR_STATE;
    }
    if (stride_floats < 11 + ctx->cfg.n_extra_dims || (n && (!coords_dev || !out_dev))) {
        set_error("network_inference: bad stride or null pointer");
        return INSTA_ERR_INVALID;
    }
    if (insta::mfma_network_supported(ctx->cfg) && ctx->cfg.n_extra_dims == 0)
        return insta::launch_network_mfma_aos(ctx, coords_dev, n, stride_floats, out_dev, (hipStream_t)stream);
    // per-sample extra values (NerfCoordinate tail) take the FP32 per-sample kernel at any precision
    if (!insta::network_supported(ctx->cfg) || (!insta::scalar_precision(ctx->cfg) && ctx->cfg.n_extra_dims == 0)) {
        set_error("network_inference: configuration not supported by the fused MLP kernels");
        return INSTA_ERR_UNSUPPORTED;
    }
    return insta::launch_network_aos(ctx, coords_dev, n, stride_floats, out_dev, (hipStream_t)stream);
}

insta_status insta_render(insta_ctx* ctx, const insta_camera* cam, uint32_t width, uint32_t height, int32_t frame,
                          int32_t occupancy, uint32_t spp, float* rgba_dev, float* depth_dev, uint32_t* steps_dev,
                          void* stream) {
    INSTA_CHECK_CTX(ctx);
    const insta_status st = insta::check_render_args(ctx, cam, width, height, frame, occupancy, rgba_dev);
    if (st != INSTA_OK) return st;
    return insta::render(ctx, ctx->ws, cam, width, height, frame, occupancy, spp, rgba_dev, depth_dev, steps_dev,
                         (hipStream_t)stream);
}

namespace {
bool frame_config_ok(const insta_frame_config* fc) {
    return fc && fc->tonemap_curve <= INSTA_TONEMAP_REINHARD && std::isfinite(fc->exposure) &&
           std::isfinite(fc->background[0]) && std::isfinite(fc->background[1]) &&
           std::isfinite(fc->background[2]) && std::isfinite(fc->background[3]);
}
}  // namespace

insta_status insta_network_forward(insta_ctx* ctx, const float* coords_dev, uint32_t n, uint32_t stride_floats,
                                   float* out_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!coords_dev || !out_dev || n == 0 || stride_floats != 11 + ctx->cfg.n_extra_dims) {
        set_error("network_forward: null pointer, empty batch or stride other than 11 + n_extra_dims");
        return INSTA_ERR_INVALID;
    }
    if (!ctx->params_set) {
        set_error("network_forward: parameters not set");
        return INSTA_ERR_STATE;
    }
    if (!insta::train_supported(ctx->cfg)) {
        set_error("network_forward: network shape not supported by the training kernels");
        return INSTA_ERR_UNSUPPORTED;
    }
    return insta::network_forward(ctx, coords_dev, n, stride_floats, out_dev, (hipStream_t)stream);
}

insta_status insta_network_backward(insta_ctx* ctx, const float* coords_dev, uint32_t n, uint32_t stride_floats,
                                    const float* dL_doutput_dev, float* grad_params_dev, int accumulate,
                                    float* dL_dinput_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!coords_dev || !dL_doutput_dev || !grad_params_dev || stride_floats != 11 + ctx->cfg.n_extra_dims) {
        set_error("network_backward: null pointer or stride other than 11 + n_extra_dims");
        return INSTA_ERR_INVALID;
    }
    if (n == 0 || insta::network_forward_batch(ctx) != n) {
        set_error("network_backward: no stored forward of this batch size (insta_network_forward first)");
        return INSTA_ERR_STATE;
    }
    return insta::network_backward(ctx, n, stride_floats, dL_doutput_dev, grad_params_dev, accumulate != 0,
                                   dL_dinput_dev, (hipStream_t)stream);
}

insta_status insta_network_density(insta_ctx* ctx, const float* positions_dev, uint32_t n, float* out_dev,
                                   void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!positions_dev || !out_dev) {
        set_error("network_density: null pointer");
        return INSTA_ERR_INVALID;
    }
    if (!ctx->params_set) {
        set_error("network_density: parameters not set");
        return INSTA_ERR_STATE;
    }
    if (!insta::train_supported(ctx->cfg)) {
        set_error("network_density: network shape not supported");
        return INSTA_ERR_UNSUPPORTED;
    }
    return insta::launch_density_outputs(ctx, positions_dev, n, out_dev, (hipStream_t)stream);
}

insta_status insta_frame_buffer(const float* rgba_dev, const float* depth_dev, uint32_t n, uint32_t spp,
                                const insta_frame_config* fc, float* accum_dev, float* out_dev,
                                float* depth_out_dev, void* stream) {
    if (!rgba_dev || !accum_dev || !out_dev || !frame_config_ok(fc)) {
        set_error("frame_buffer: null buffer or bad frame configuration");
        return INSTA_ERR_INVALID;
    }
    return insta::launch_frame_buffer(reinterpret_cast<const float4*>(rgba_dev), depth_dev, n, spp, *fc,
                                      reinterpret_cast<float4*>(accum_dev), reinterpret_cast<float4*>(out_dev),
                                      depth_out_dev, (hipStream_t)stream);
}

insta_status insta_render_frame(insta_ctx* ctx, const insta_camera* cam, uint32_t width, uint32_t height,
                                int32_t frame, int32_t occupancy, uint32_t spp, const insta_frame_config* fc,
                                float* accum_dev, float* out_dev, float* depth_out_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!accum_dev || !out_dev || !frame_config_ok(fc)) {
        set_error("render_frame: null buffer or bad frame configuration");
        return INSTA_ERR_INVALID;
    }
    // the composited image goes through out_dev (shaded in place), the
    // composited depth through depth_out_dev
    insta_status st = insta_render(ctx, cam, width, height, frame, occupancy, spp, out_dev, depth_out_dev, nullptr,
                                   stream);
    if (st != INSTA_OK) return st;
    return insta::launch_frame_buffer(reinterpret_cast<const float4*>(out_dev), depth_out_dev, width * height, spp,
                                      *fc, reinterpret_cast<float4*>(accum_dev), reinterpret_cast<float4*>(out_dev),
                                      depth_out_dev, (hipStream_t)stream);
}

insta_status insta_get_deform_misses(insta_ctx* ctx, uint64_t* misses, int reset, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!misses) {
        set_error("get_deform_misses: null pointer");
        return INSTA_ERR_INVALID;
    }
    const hipStream_t s = (hipStream_t)stream;
    uint32_t h = 0;
    INSTA_HIP(hipMemcpyAsync(&h, ctx->d_misses, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    if (reset) INSTA_HIP(hipMemsetAsync(ctx->d_misses, 0, sizeof(uint32_t), s));
    INSTA_HIP(hipStreamSynchronize(s));
    *misses = h;
    return INSTA_OK;
}

insta_status insta_get_render_stats(insta_ctx* ctx, insta_render_stats* out) {
    INSTA_CHECK_CTX(ctx);
    if (!out) {
        set_error("get_render_stats: null pointer");
        return INSTA_ERR_INVALID;
    }
    *out = ctx->ws.stats;
    return INSTA_OK;
}

insta_status insta_set_profiling(insta_ctx* ctx, int enable) {
    INSTA_CHECK_CTX(ctx);
    if (enable < 0 || enable > 2) {
        set_error("set_profiling: enable is 0, 1 or 2");
        return INSTA_ERR_INVALID;
    }
    return insta::set_profiling(ctx, enable);
}

insta_status insta_get_profile(insta_ctx* ctx, insta_profile* out) {
    INSTA_CHECK_CTX(ctx);
    if (!out) {
        set_error("get_profile: null pointer");
        return INSTA_ERR_INVALID;
    }
    *out = ctx->prof;
    if (ctx->d_trav_counts) {
        // renders and training steps synchronise their stream before they
        // return, so the counts are final: read them on the context's own
        // non-blocking stream rather than stalling the device
        unsigned long long c[insta::kTravCounters];
        if (!ctx->aux_stream) INSTA_HIP(hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
        INSTA_HIP(hipMemcpyAsync(c, ctx->d_trav_counts, sizeof(c), hipMemcpyDeviceToHost, ctx->aux_stream));
        INSTA_HIP(hipStreamSynchronize(ctx->aux_stream));
        out->bvh_queries = c[0];
        out->bvh_inner = c[1];
        out->bvh_leaves = c[2];
        out->bvh_triangles = c[3];
        out->list_queries = c[4];
        out->list_entries = c[5];
    }
    return INSTA_OK;
}

}  // extern "C"

void insta_default_frame_config(insta_frame_config* fc) {
    if (!fc) return;
    *fc = insta_frame_config{};
    fc->exposure = 0.0f;
    for (int c = 0; c < 4; ++c) fc->background[c] = 1.0f;
    fc->tonemap_curve = INSTA_TONEMAP_IDENTITY;
    fc->output_srgb = 1;
    fc->linear_colors = 0;
}

void insta_default_train_config(insta_train_config* tc) {
    if (!tc) return;
    tc->n_rays = 0;
    tc->target_samples = 1u << 18;
    tc->max_samples = 1u << 22;
    tc->learning_rate = 1e-2f;
    tc->beta1 = 0.9f;
    tc->beta2 = 0.99f;
    tc->epsilon = 1e-15f;
    tc->loss_scale = 1.0f;
    tc->grid_grad_fp16 = 0;
    tc->apply_update = 1;
    tc->seed = 1337u;
    tc->occupancy_slot = INSTA_OCCUPANCY_PER_FRAME;
    tc->l2_reg = 0.0f;
    tc->ema_decay = 0.0f;
    tc->seg_weight = 1.0f;     // seg_mask_supervision_lambda (testbed.h:623)
    tc->depth_weight = 1.5f;   // depth_supervision_lambda (testbed.h:622)
}

insta_status insta_set_training_set(insta_ctx* ctx, uint32_t n_images, uint32_t width, uint32_t height,
                                    const float* rgba_dev, const insta_camera* cameras, const int32_t* frames,
                                    void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (n_images == 0 || width == 0 || height == 0 || !rgba_dev || !cameras || !frames ||
        (uint64_t)n_images * width * height > (1ull << 32)) {
        set_error("set_training_set: bad sizes or null pointer");
        return INSTA_ERR_INVALID;
    }
    for (uint32_t i = 0; i < n_images; ++i) {
        const int32_t f = frames[i];
        if (f < 0 || f >= (int32_t)ctx->cfg.max_frames || !ctx->slot_built[f + 1]) {
            set_error("set_training_set: an image refers to a frame whose mesh is not built");
            return INSTA_ERR_STATE;
        }
        if (!(cameras[i].focal_px[0] > 0.0f) || !(cameras[i].focal_px[1] > 0.0f)) {
            set_error("set_training_set: focal lengths must be positive");
            return INSTA_ERR_INVALID;
        }
    }
    if (!insta::train_supported(ctx->cfg)) {
        set_error("set_training_set: training supports width-64 MLPs with 1 density / 2 colour hidden layers, "
                  "F = 2, L in {8, 16}, no extra dims");
        return INSTA_ERR_UNSUPPORTED;
    }
    return insta::set_training_set(ctx, n_images, width, height, rgba_dev, cameras, frames, (hipStream_t)stream);
}

insta_status insta_set_training_supervision(insta_ctx* ctx, const uint8_t* seg_dev, const float* depth_dev,
                                            void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!insta::train_has_set(ctx)) {
        set_error("set_training_supervision: no training set");
        return INSTA_ERR_STATE;
    }
    return insta::set_training_supervision(ctx, seg_dev, depth_dev, (hipStream_t)stream);
}

insta_status insta_get_inference_params(insta_ctx* ctx, float* params_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!params_dev || !ctx->params_set) {
        set_error("get_inference_params: null pointer or parameters not set");
        return !params_dev ? INSTA_ERR_INVALID : INSTA_ERR_STATE;
    }
    INSTA_HIP(hipMemcpyAsync(params_dev, insta::net_params(ctx), ctx->layout.n_params * sizeof(float),
                             hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return INSTA_OK;
}

insta_status insta_reserve(insta_ctx* ctx, const insta_reserve_config* r, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!r || (uint64_t)r->render_pixels > (1ull << 26) || r->train_rays > (1u << 24) ||
        r->train_samples > (1u << 26) || r->train_batch > (1u << 24)) {
        set_error("reserve: null pointer or size beyond the limits of render and train_step");
        return INSTA_ERR_INVALID;
    }
    if (r->render_pixels) {
        const insta_status st = insta::reserve_render_ws(ctx->ws, r->render_pixels);
        if (st != INSTA_OK) return st;
    }
    if (r->train_rays || r->train_samples || r->train_batch) {
        const insta_status st =
            insta::reserve_training(ctx, r->train_rays, r->train_samples, r->train_batch, (hipStream_t)stream);
        if (st != INSTA_OK) return st;
    }
    return INSTA_OK;
}

insta_status insta_train_step(insta_ctx* ctx, const insta_train_config* tc, insta_train_stats* stats, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!tc || tc->max_samples == 0 || tc->max_samples > (1u << 26) || tc->n_rays > (1u << 24) ||
        tc->target_samples == 0 || tc->target_samples > (1u << 24) || !(tc->loss_scale > 0.0f) ||
        !(tc->learning_rate >= 0.0f) || !(tc->l2_reg >= 0.0f) || !(tc->ema_decay >= 0.0f && tc->ema_decay < 1.0f) ||
        !(tc->seg_weight >= 0.0f && tc->seg_weight <= 1e6f) || !(tc->depth_weight >= 0.0f && tc->depth_weight <= 1e6f)) {
        set_error("train_step: bad training configuration");
        return INSTA_ERR_INVALID;
    }
    if (!insta::train_has_set(ctx)) {
        set_error("train_step: no training set");
        return INSTA_ERR_STATE;
    }
    if (!ctx->params_set) {
        set_error("train_step: parameters not set");
        return INSTA_ERR_STATE;
    }
    if (tc->occupancy_slot != INSTA_OCCUPANCY_PER_FRAME) {
        if (tc->occupancy_slot < -1 || tc->occupancy_slot >= (int32_t)ctx->cfg.max_frames) {
            set_error("train_step: bad occupancy slot");
            return INSTA_ERR_INVALID;
        }
        if (!ctx->grid_built[tc->occupancy_slot + 1]) {
            set_error("train_step: the occupancy slot has no grid");
            return INSTA_ERR_STATE;
        }
    } else {
        for (const int32_t f : insta::train_frames(ctx)) {
            if (!ctx->slot_built[f + 1] || !ctx->grid_built[f + 1]) {
                set_error("train_step: frame " + std::to_string(f) +
                          " of the training set has no mesh or occupancy grid (rays march their frame's grid)");
                return INSTA_ERR_STATE;
            }
        }
    }
    return insta::train_step(ctx, *tc, stats, (hipStream_t)stream);
}

insta_status insta_train_loop(insta_ctx* ctx, const insta_train_config* tc, uint32_t n_steps,
                              const insta_density_grid_config* dg, int32_t density_slot, uint32_t n_frames,
                              insta_train_stats* last, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!tc || !dg) {
        set_error("train_loop: null configuration");
        return INSTA_ERR_INVALID;
    }
    const uint64_t g3 = (uint64_t)ctx->cfg.grid_size * ctx->cfg.grid_size * ctx->cfg.grid_size;
    insta_train_config step_cfg = *tc;
    step_cfg.occupancy_slot = density_slot;
    for (uint32_t k = 0; k < n_steps; ++k) {
        // the optimizer step count plays the reference's m_training_step
        const uint32_t s = ctx->train ? insta::train_step_count(ctx) : 0u;
        const uint32_t every = std::min(std::max(s / 16u, 1u), 16u);
        if (s % every == 0) {
            insta_density_grid_config c = *dg;
            c.n_uniform = (uint32_t)(s < 256 ? g3 : g3 / 4);
            c.n_nonuniform = (uint32_t)(s < 256 ? 0 : g3 / 4);
            insta_status st = insta_update_density_grid(ctx, density_slot, n_frames, &c, nullptr, nullptr, stream);
            if (st != INSTA_OK) return st;
        }
        insta_status st = insta_train_step(ctx, &step_cfg, last, stream);
        if (st != INSTA_OK) return st;
    }
    return INSTA_OK;
}

insta_status insta_get_gradients(insta_ctx* ctx, float* grads_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!grads_dev) {
        set_error("get_gradients: null pointer");
        return INSTA_ERR_INVALID;
    }
    if (!insta::train_has_gradients(ctx)) {
        set_error("get_gradients: no training step has run");
        return INSTA_ERR_STATE;
    }
    return insta::get_gradients(ctx, grads_dev, (hipStream_t)stream);
}

insta_status insta_get_params(insta_ctx* ctx, float* params_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!params_dev) {
        set_error("get_params: null pointer");
        return INSTA_ERR_INVALID;
    }
    if (!ctx->params_set) {
        set_error("get_params: parameters not set");
        return INSTA_ERR_STATE;
    }
    INSTA_HIP(hipMemcpyAsync(params_dev, ctx->d_params, ctx->layout.n_params * sizeof(float),
                             hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return INSTA_OK;
}

insta_status insta_reset_optimizer(insta_ctx* ctx, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!ctx->train) return INSTA_OK;
    return insta::reset_optimizer(ctx, (hipStream_t)stream);
}

insta_status insta_deform_positions(insta_ctx* ctx, float* positions_dev, uint32_t n, uint32_t n_frames, uint32_t seed,
                                    uint64_t rng_offset, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if ((n && !positions_dev) || n_frames == 0 || n_frames > ctx->cfg.max_frames) {
        set_error("deform_positions: null pointer or bad frame count");
        return INSTA_ERR_INVALID;
    }
    // frames drawn: 0 .. max(n_frames - 2, 0)
    for (uint32_t f = 0; f < std::max(n_frames - 1, 1u); ++f)
        if (!ctx->slot_built[0] || !ctx->slot_built[f + 1]) {
            set_error("deform_positions: a frame in 0 .. n_frames - 2 is not built");
            return INSTA_ERR_STATE;
        }
    insta::Pcg32 rng = insta::Pcg32::seeded(seed);
    rng.advance(rng_offset);
    return insta::launch_deform_positions(ctx, positions_dev, n, n_frames, rng, (hipStream_t)stream);
}

insta_status insta_update_density_grid(insta_ctx* ctx, int32_t slot, uint32_t n_frames,
                                       const insta_density_grid_config* cfg, float* grid_dev, float* mean,
                                       void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (!cfg || slot < -1 || slot >= (int32_t)ctx->cfg.max_frames || n_frames == 0 ||
        n_frames > ctx->cfg.max_frames || (uint64_t)cfg->n_uniform + cfg->n_nonuniform == 0 ||
        (uint64_t)cfg->n_uniform + cfg->n_nonuniform > (1ull << 26) || !(cfg->decay >= 0.0f && cfg->decay <= 1.0f)) {
        set_error("update_density_grid: bad slot, frame count or configuration");
        return INSTA_ERR_INVALID;
    }
    if (!ctx->params_set) {
        set_error("update_density_grid: parameters not set");
        return INSTA_ERR_STATE;
    }
    if (!insta::train_supported(ctx->cfg)) {
        set_error("update_density_grid: network shape not supported (see insta_set_training_set)");
        return INSTA_ERR_UNSUPPORTED;
    }
    for (uint32_t f = 0; f < std::max(n_frames - 1, 1u); ++f)
        if (!ctx->slot_built[f + 1]) {
            set_error("update_density_grid: a frame in 0 .. n_frames - 2 is not built");
            return INSTA_ERR_STATE;
        }
    return insta::update_density_grid(ctx, (uint32_t)(slot + 1), n_frames, *cfg, grid_dev, mean,
                                      (hipStream_t)stream);
}

insta_status insta_reset_density_grid(insta_ctx* ctx) {
    INSTA_CHECK_CTX(ctx);
    return insta::reset_density_grid(ctx);
}

insta_status insta_get_occupancy(insta_ctx* ctx, int32_t slot, uint8_t* bitfield_dev, void* stream) {
    INSTA_CHECK_CTX(ctx);
    if (slot < -1 || slot >= (int32_t)ctx->cfg.max_frames || !bitfield_dev) {
        set_error("get_occupancy: bad slot or null pointer");
        return INSTA_ERR_INVALID;
    }
    if (!ctx->grid_built[slot + 1]) {
        set_error("get_occupancy: slot has no occupancy grid");
        return INSTA_ERR_STATE;
    }
    const size_t G = ctx->cfg.grid_size;
    INSTA_HIP(hipMemcpyAsync(bitfield_dev, ctx->grid((uint32_t)(slot + 1)), G * G * G / 8, hipMemcpyDeviceToDevice,
                             (hipStream_t)stream));
    return INSTA_OK;
}
