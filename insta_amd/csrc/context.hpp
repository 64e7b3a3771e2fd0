// context.hpp -- host-side state of one avatar context and the launchers that
// each kernel translation unit provides.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.hpp"

namespace insta {
struct TrainState;
struct DensityGridState;

// Buffers of one render in flight: per pixel ray state, per sample the
// march/deform/network stages, and the round counters.
struct RenderWS {
    uint32_t pixels = 0;            // capacity in pixels
    uint64_t samples = 0;           // sample-buffer capacity
    uint32_t samples_per_pixel = 128;  // sample-buffer sizing (more rounds when exceeded)
    float4* r_o = nullptr;          // origin xyz, t
    float4* r_d = nullptr;          // dir xyz, max_weight
    float4* r_rgba = nullptr;
    float* r_depth = nullptr;
    uint32_t* r_steps = nullptr;    // samples composited
    uint32_t* r_flags = nullptr;    // bit1: exited the box in the last march
    uint32_t* alive[2] = {nullptr, nullptr};
    uint32_t* r_off = nullptr;      // per alive-list entry: first sample
    uint32_t* r_cnt = nullptr;      // per alive-list entry: samples in this round
    float4* s_pos = nullptr;        // posed xyz, ray id bits
    float4* s_cpos = nullptr;       // canonical xyz
    float4* s_cdir = nullptr;       // canonical direction (warped to [0,1])
    float4* s_out = nullptr;        // rgb logits, density logit
    int32_t* s_hit = nullptr;       // per sample slot: march t values, then nearest triangle (BVH order)
    uint32_t* d_counters = nullptr; // [16]: alive counts, sample counter, stats
    uint32_t* h_counters = nullptr; // pinned mirror
    insta_render_stats stats{};
    uint64_t bytes = 0;             // device bytes held
    bool external = false;          // buffers carved from caller memory (bind_render_ws)
};
// words of a render's round counters (RenderWS::d_counters)
constexpr uint32_t kRenderCounterWords = 16;
// device bytes a workspace for n_pixels needs (with spp samples per pixel,
// or an explicit sample capacity per round)
uint64_t render_ws_bytes(uint32_t n_pixels, uint32_t samples_per_pixel);
uint64_t render_ws_bytes_cap(uint32_t n_pixels, uint64_t sample_cap);
void bind_render_ws_cap(RenderWS& ws, char* base, uint32_t n_pixels, uint64_t sample_cap);
void release_render_ws(RenderWS& ws);
// allocate ws (and its counters) for n_pixels (insta_reserve)
insta_status reserve_render_ws(RenderWS& ws, uint32_t n_pixels);
// point ws's buffers into [base, base + render_ws_bytes(n_pixels, spp)) (no
// allocation; the counters must already exist)
void bind_render_ws(RenderWS& ws, char* base, uint32_t n_pixels, uint32_t samples_per_pixel);
}

struct insta_ctx {
    insta_config cfg{};
    insta_layout layout{};
    int device = 0;

    // topology (shared by all slots)
    uint32_t nv = 0, nf = 0;
    uint32_t bvh_k = 0, leaf_base = 0, n_nodes = 0;
    uint32_t node_cap = 0;          // nodes reserved per slot (for max_triangles)
    uint32_t n_slots = 0;           // max_frames + 1
    std::vector<uint8_t> slot_built;
    std::vector<uint8_t> grid_built;

    int32_t* d_faces = nullptr;     // [nf*3]
    uint32_t* d_masks = nullptr;    // [nf]
    int4* d_adj = nullptr;          // [nf] (3 edge neighbours, -1 padded)
    float* d_verts = nullptr;       // [n_slots][max_vertices*3]
    float4* d_nodes = nullptr;      // [n_slots][node_cap*2]
    float4* d_btris = nullptr;      // [n_slots][max_triangles*3]
    float4* d_proj = nullptr;       // [n_slots][max_triangles*3]
    float4* d_proj_inv = nullptr;   // [n_slots][max_triangles*3]
    float4* d_cen = nullptr;        // [n_slots][max_triangles]
    uint8_t* d_grids = nullptr;     // [n_slots][G^3/8] Morton-order bitfields (the reference layout)
    uint8_t* d_grids_lin = nullptr; // [n_slots][G^3/8] same bits, linear x + G(y + G z) order (march)
    uint8_t* d_built = nullptr;     // [n_slots] 1 once a slot's mesh is built
    uint16_t* d_build_lists = nullptr;  // [n_slots][2][3][max_triangles] BVH build scratch
    uint32_t* d_build_q = nullptr;      // [n_slots][3][max_triangles]
    float render_lo[3] = {0, 0, 0}, render_hi[3] = {1, 1, 1};

    // nearest-triangle candidate lists per slot (cfg.cell_split > 0; layout
    // in common.hpp), sized at creation from cell_list_cells / _entries
    uint64_t* d_cl_bits = nullptr;  // [n_slots][G^3/64]
    uint32_t* d_cl_rank = nullptr;  // [n_slots][G^3/64]
    uint32_t* d_cl_cells = nullptr; // [n_slots][cell_list_cells] occupied cells (linear index)
    uint2* d_cl_head = nullptr;     // [n_slots][cell_list_cells * S^3]
    uint32_t* d_cl_ent = nullptr;   // [n_slots][cell_list_entries]
    uint32_t* d_cl_state = nullptr; // [n_slots][kClState]

    // parameters
    float* d_params = nullptr;      // FP32 master copy [n_params] (the weights training updates)
    const float* net_src = nullptr; // the parameters the inference network is prepared from:
                                    // d_params, or the debiased weight average under an EMA optimizer
    void* d_wfrag = nullptr;        // MLP weights arranged as MFMA A fragments
    void* d_table_h = nullptr;      // FP16 copy of the hash tables (reduced precision)
    float* d_frame_extra = nullptr; // [max_frames][n_extra_dims] per-frame extra dims (n_extra_dims > 0)
    float* d_extra_bias = nullptr;  // [n_slots][64] W0[:, enc:] . x: row 0 x = 1, row 1 + f frame f's values
    bool params_set = false;

    // render workspace of insta_render (grown on demand); the server keeps
    // one per slot
    insta::RenderWS ws;

    // training state (training set, workspace, gradients, Adam moments)
    insta::TrainState* train = nullptr;
    // density-grid EMA state (insta_update_density_grid)
    insta::DensityGridState* density = nullptr;

    // optional per-kernel-class event timing of insta_render
    bool profiling = false;
    unsigned long long* d_trav_counts = nullptr;  // [kTravCounters] while traversal counting is on
    hipStream_t aux_stream = nullptr;             // non-blocking stream for host reads of counters
    hipEvent_t prof_ev[2 * INSTA_KERNEL_CLASSES] = {};
    bool prof_pending[INSTA_KERNEL_CLASSES] = {};
    uint64_t prof_pending_items[INSTA_KERNEL_CLASSES] = {};
    insta_profile prof{};

    // samples the AoS deformation entry points left undeformed because the
    // nearest-triangle query found nothing (src/core.cu:482-485)
    uint32_t* d_misses = nullptr;

    // misc scratch for AoS entry points
    void* d_scratch = nullptr;
    size_t scratch_bytes = 0;

    insta_ctx() = default;
    insta_ctx(const insta_ctx&) = delete;
    insta_ctx& operator=(const insta_ctx&) = delete;

    insta::MeshView mesh(uint32_t slot) const {
        insta::MeshView m;
        const size_t T = cfg.max_triangles;
        m.nodes = d_nodes + (size_t)slot * node_cap * 2;
        m.btris = d_btris + (size_t)slot * T * 3;
        m.proj = d_proj + (size_t)slot * T * 3;
        m.proj_inv = d_proj_inv + (size_t)slot * T * 3;
        m.cen = d_cen + (size_t)slot * T;
        m.nf = nf;
        m.leaf_base = leaf_base;
        m.k = bvh_k;
        return m;
    }
    uint8_t* grid(uint32_t slot) const {
        const size_t G = cfg.grid_size;
        return d_grids + (size_t)slot * (G * G * G / 8);
    }
    uint8_t* grid_lin(uint32_t slot) const {
        const size_t G = cfg.grid_size;
        return d_grids_lin + (size_t)slot * (G * G * G / 8);
    }
    // candidate lists of a slot (all null when cell_split is 0)
    insta::CellListView cells(uint32_t slot) const {
        insta::CellListView v{};
        if (!d_cl_state) return v;
        const size_t G = cfg.grid_size, S = cfg.cell_split, words = G * G * G / 64;
        v.bits = d_cl_bits + slot * words;
        v.rank = d_cl_rank + slot * words;
        v.head = d_cl_head + slot * (size_t)cfg.cell_list_cells * S * S * S;
        v.ent = d_cl_ent + slot * (size_t)cfg.cell_list_entries;
        v.state = d_cl_state + insta::kClState * (size_t)slot;
        v.G = cfg.grid_size;
        v.S = cfg.cell_split;
        v.cap = cfg.cell_list_cells;
        return v;
    }
};

namespace insta {

// the parameters the inference network reads (see insta_ctx::net_src)
inline const float* net_params(const insta_ctx* c) { return c->net_src ? c->net_src : c->d_params; }

template <class T>
inline void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void set_error(const std::string& msg);
insta_status load_config_json(const char* path, insta_config* cfg, insta_run_config* run);
insta_status hip_fail(hipError_t e, const char* what);

#define INSTA_HIP(call)                                              \
    do {                                                             \
        hipError_t _e = (call);                                      \
        if (_e != hipSuccess) return insta::hip_fail(_e, #call);     \
    } while (0)

// Hash-grid level constants passed by value to kernels.
struct GridParams {
    float scale[INSTA_MAX_LEVELS];
    uint32_t res[INSTA_MAX_LEVELS];
    uint32_t offset[INSTA_MAX_LEVELS];
    uint32_t size[INSTA_MAX_LEVELS];
    uint32_t dense_mask;
};
GridParams grid_params(const insta_ctx* ctx);

// launchers (one per translation unit)
insta_status launch_build_slots(insta_ctx* ctx, uint32_t first_slot, uint32_t n_slots, const float* verts,
                                hipStream_t stream);
insta_status launch_blend_expression(insta_ctx* ctx, uint32_t slot, const float* bases, uint32_t n_bases,
                                     const float* coeffs, hipStream_t stream);
insta_status launch_closest(insta_ctx* ctx, uint32_t slot, const float* pts, uint32_t n, int32_t* ids, float* dist,
                            bool use_lists, hipStream_t stream);
insta_status launch_deform_aos(insta_ctx* ctx, float* coords, uint32_t n, uint32_t stride, int32_t frame,
                               hipStream_t stream);
insta_status launch_deform_soa(insta_ctx* ctx, RenderWS& ws, int32_t frame, const uint32_t* n_dev, uint32_t n_max,
                               uint32_t* miss_counter, hipStream_t stream);
// positions [n][3] in place, each through a random frame of 0..n_frames-2
insta_status launch_deform_positions(insta_ctx* ctx, float* pos, uint32_t n, uint32_t n_frames, const Pcg32& rng,
                                     hipStream_t stream);
// samples s_pos (xyz, ray id bits) with ray directions r_d; frame per ray
// from ray_frame when non-null
// miss_counter (nullable) counts in-box samples whose query found no triangle
insta_status launch_deform_samples(insta_ctx* ctx, int32_t frame, const int32_t* ray_frame, const uint32_t* n_dev,
                                   uint32_t n_max, const float4* s_pos, const float4* r_d, int32_t* s_hit,
                                   float4* s_cpos, float4* s_cdir, uint32_t* miss_counter, hipStream_t stream);
insta_status launch_occupancy(insta_ctx* ctx, uint32_t slot, float threshold, hipStream_t stream);
// candidate lists (celllist.hip): device buffers, the build over a slot's
// current linear occupancy bits, and the invalidation when its mesh changes
insta_status alloc_cell_lists(insta_ctx* ctx);
void free_cell_lists(insta_ctx* ctx);
insta_status launch_cell_lists(insta_ctx* ctx, uint32_t slot, hipStream_t stream);
insta_status launch_cell_lists_invalidate(insta_ctx* ctx, uint32_t first, uint32_t n, hipStream_t stream);
// occupancy of the slot's posed mesh, its linear copy and its candidate lists
insta_status launch_mesh_occupancy(insta_ctx* ctx, uint32_t slot, float threshold, hipStream_t stream);
insta_status launch_grid_linear(insta_ctx* ctx, uint32_t slot, hipStream_t stream);
insta_status launch_set_built(insta_ctx* ctx, uint32_t first, uint32_t n, hipStream_t stream);
// render pipeline (render.hip) on an explicit workspace
insta_status render(insta_ctx* ctx, RenderWS& ws, const insta_camera* cam, uint32_t width, uint32_t height,
                    int32_t frame, int32_t occupancy, uint32_t spp, float* rgba, float* depth, uint32_t* steps,
                    hipStream_t stream);
// The render split at its one host synchronisation: enqueue_render queues ray
// init, `rounds` march/deform/network/composite rounds (live counts stay on
// the device) and the final image write, with no host synchronisation, so a
// stream capture of it is a complete graph; finish_render reads the counters,
// runs any further rounds the rays still need and rewrites the image.
insta_status enqueue_render(insta_ctx* ctx, RenderWS& ws, const insta_camera* cam, uint32_t width, uint32_t height,
                            int32_t frame, int32_t occupancy, uint32_t spp, uint32_t rounds, float* rgba,
                            float* depth, uint32_t* steps, hipStream_t stream);
insta_status finish_render(insta_ctx* ctx, RenderWS& ws, const insta_camera* cam, uint32_t width, uint32_t height,
                           int32_t frame, int32_t occupancy, uint32_t spp, uint32_t rounds, float* rgba, float* depth,
                           uint32_t* steps, hipStream_t stream);
void free_workspace(insta_ctx* ctx);
// render buffer stages (shade, accumulate, tonemap) on a composited image
insta_status launch_frame_buffer(const float4* rgba, const float* depth, uint32_t n, uint32_t spp,
                                 const insta_frame_config& fc, float4* accum, float4* out, float* depth_out,
                                 hipStream_t stream);
insta_status set_profiling(insta_ctx* ctx, int level);
// resolve this device's fixed render scratch (noalloc.hip); called at set-up
insta_status noalloc_prepare_device(int device);
// the argument and state checks of insta_render
insta_status check_render_args(const insta_ctx* ctx, const insta_camera* cam, uint32_t width, uint32_t height,
                               int32_t frame, int32_t occupancy, const float* rgba);
// training (train.hip)
bool train_supported(const insta_config& cfg);
void free_train(insta_ctx* ctx);
insta_status set_training_set(insta_ctx* ctx, uint32_t n_images, uint32_t width, uint32_t height, const float* rgba,
                              const insta_camera* cams, const int32_t* frames, hipStream_t stream);
insta_status train_step(insta_ctx* ctx, const insta_train_config& tc, insta_train_stats* stats, hipStream_t stream);
// training state and workspace for up to `rays` rays, `samples` marched
// samples and a compacted batch of `batch` (insta_reserve)
insta_status reserve_training(insta_ctx* ctx, uint32_t rays, uint32_t samples, uint32_t batch, hipStream_t stream);
bool train_has_gradients(const insta_ctx* ctx);
// new weights: inference reads them directly, the Ema average restarts
void reset_inference_average(insta_ctx* ctx);
// density grid (density.hip)
void free_density_grid(insta_ctx* ctx);
insta_status update_density_grid(insta_ctx* ctx, uint32_t slot, uint32_t n_frames, const insta_density_grid_config& c,
                                 float* grid_out, float* mean_out, hipStream_t stream);
insta_status reset_density_grid(insta_ctx* ctx);
bool train_has_set(const insta_ctx* ctx);
// one training image's camera on the device: 3x4 matrix (columns right, down,
// forward, position; row-major), focal lengths and principal point in
// pixels / image fractions, its posed frame
struct TrainImage {
    float M[12];
    float fx, fy, cx, cy;
    int32_t frame;
    uint32_t pad[3];
};
struct TrainSetView {
    const TrainImage* meta = nullptr;  // device [n_images]
    uint32_t n_images = 0, width = 0, height = 0;
};
TrainSetView train_set_view(const insta_ctx* ctx);
const std::vector<int32_t>& train_frames(const insta_ctx* ctx);  // frame of each training image
insta_status set_training_supervision(insta_ctx* ctx, const uint8_t* seg, const float* depth, hipStream_t stream);
uint32_t train_step_count(const insta_ctx* ctx);  // training steps taken (the reference's m_training_step)
// snapshot state (snapshot.hip): host copies of the training and density-grid state
struct TrainSnapshot {
    bool present = false;
    uint32_t step = 0, rays_per_batch = 0;  // step: training steps taken (snapshot "training_step")
    uint32_t adam_step = 0;                 // Adam's step (snapshot optimizer "step"), with m and v
    uint64_t rays_total = 0;
    std::vector<float> m, v;  // Adam moments, empty unless saved
    // training set cameras: per image the 3x4 matrix, focal (px), principal point
    uint32_t n_images = 0, width = 0, height = 0;
    std::vector<float> cams;  // [n_images][16]: matrix[12], fx, fy, cx, cy
};
struct DensitySnapshot {
    std::vector<float> grid;  // [G^3] Morton order, empty if never populated
    uint32_t step = 0;
    uint64_t rng_state = 0, rng_inc = 0;
};
insta_status train_export(insta_ctx* ctx, bool with_optimizer, TrainSnapshot& out, hipStream_t stream);
insta_status train_import(insta_ctx* ctx, const TrainSnapshot& in, hipStream_t stream);
insta_status density_export(insta_ctx* ctx, DensitySnapshot& out, hipStream_t stream);
// grid, then its mean and bitfield into occupancy slot `slot` (0 canonical, f + 1 frame f)
insta_status density_import(insta_ctx* ctx, const DensitySnapshot& in, uint32_t slot, hipStream_t stream);
insta_status get_gradients(insta_ctx* ctx, float* out, hipStream_t stream);
// network plugin (train.hip, density.hip)
insta_status network_forward(insta_ctx* ctx, const float* coords, uint32_t n, uint32_t stride, float* out,
                             hipStream_t stream);
uint32_t network_forward_batch(const insta_ctx* ctx);  // samples of the stored forward (0: none)
insta_status network_backward(insta_ctx* ctx, uint32_t n, uint32_t stride, const float* dL_dout, float* grad_out,
                              bool accumulate, float* dL_dinput, hipStream_t stream);
insta_status launch_density_outputs(insta_ctx* ctx, const float* pos, uint32_t n, float* out, hipStream_t stream);
insta_status reset_optimizer(insta_ctx* ctx, hipStream_t stream);
insta_status launch_encode(insta_ctx* ctx, const float* pos, uint32_t n, float* out, hipStream_t stream);
insta_status launch_encode_backward(insta_ctx* ctx, const float* pos, uint32_t n, const float* dL, float* grad,
                                    hipStream_t stream);
insta_status launch_network_aos(insta_ctx* ctx, const float* coords, uint32_t n, uint32_t stride, float* out,
                                hipStream_t stream);
insta_status launch_network_soa(insta_ctx* ctx, const uint32_t* n_dev, uint32_t n_max, const float4* cpos,
                                const float4* cdir, float4* out, hipStream_t stream);
insta_status launch_network_samples(insta_ctx* ctx, const uint32_t* n_dev, uint32_t n_max, const float4* cpos,
                                    const float4* cdir, float4* out, hipStream_t stream);
bool network_supported(const insta_config& cfg);
bool mfma_network_supported(const insta_config& cfg);
// FP32 arithmetic that the per-sample (scalar FMA) MLP kernel implements
inline bool scalar_precision(const insta_config& c) {
    return c.precision == INSTA_PRECISION_FP32 || c.precision == INSTA_PRECISION_FP32_SCALAR;
}
insta_status prepare_network(insta_ctx* ctx, hipStream_t stream);
// extra-dims table rows [first, first + n) from the current weights and frame values
insta_status update_extra_rows(insta_ctx* ctx, uint32_t first, uint32_t n, hipStream_t stream);
// frame's extra dims = first n_extra_dims of coeffs (device), then its table row
insta_status launch_extra_from_coeffs(insta_ctx* ctx, uint32_t frame, const float* coeffs, uint32_t n_coeffs,
                                      hipStream_t stream);
// per-sample extra-dims selector in the canonical direction's w: 0 = all 1.0,
// 1 + f = frame f's values
constexpr uint32_t kExtraOnes = 0u;
insta_status launch_network_mfma_soa(insta_ctx* ctx, const uint32_t* n_dev, uint32_t n_max, const float4* cpos,
                                     const float4* cdir, float4* out, hipStream_t stream);
insta_status launch_network_mfma_aos(insta_ctx* ctx, const float* coords, uint32_t n, uint32_t stride, float* out,
                                     hipStream_t stream);

}  // namespace insta
