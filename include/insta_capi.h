/*
 * insta_capi.h -- C ABI of the MI355X-native INSTA per-sample hot path.
 *
 * One context (`insta_ctx`) holds one avatar: its canonical mesh, the posed
 * frames with their device-built BVHs and per-triangle frames, the network
 * parameters (hash grid + density/colour MLPs), the occupancy bitfields and
 * the render workspace.  Every entry point is stream-ordered on the
 * caller-given HIP stream (passed as `void*`, a `hipStream_t`), returns an
 * `insta_status`, and never throws; `insta_last_error()` gives a message for
 * the last failure on the calling thread.  Device pointers (`*_dev`) are
 * caller-owned device memory; the context never allocates per call once its
 * workspace has grown to the largest request seen (render workspace grows on
 * demand, BVH/frame storage is sized at creation from `max_frames` and
 * `max_triangles`).
 *
 * Coordinate layout of a sample ("NerfCoordinate", 11 + n_extra_dims floats,
 * instant-ngp/include/neural-graphics-primitives/nerf.h:86-115):
 *   [0..2] position (unit-cube world coordinates)   [3] dt (warped)
 *   [4] holder = frame id (negative: skip)           [5..7] deformed position
 *   [8..10] direction mapped to [0,1]^3 ((d+1)/2)    [11..] extra dims
 */
#ifndef INSTA_CAPI_H
#define INSTA_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum insta_status {
    INSTA_OK = 0,
    INSTA_ERR_INVALID = 1,     /* bad argument or size */
    INSTA_ERR_HIP = 2,         /* a HIP runtime call failed */
    INSTA_ERR_OOM = 3,         /* device allocation failed */
    INSTA_ERR_STATE = 4,       /* call order violated (e.g. render before mesh) */
    INSTA_ERR_UNSUPPORTED = 5  /* configuration outside what the kernels handle */
} insta_status;

/* Arithmetic of the network: FP32 MFMA (v_mfma_f32_16x16x4f32), FP16 or BF16
 * MFMA with FP32 accumulation and FP16 hash tables, or FP32 per-sample FMA
 * chains without matrix cores (the plain FP32 path). */
enum {
    INSTA_PRECISION_FP32 = 0,
    INSTA_PRECISION_FP16 = 1,
    INSTA_PRECISION_BF16 = 2,
    INSTA_PRECISION_FP32_SCALAR = 3
};

/* Network/encoding/march configuration.  Field meanings follow the
 * reference's JSON keys (configs/nerf/main.json:33-69) and testbed constants
 * (instant-ngp/src/testbed_nerf.cu:48-68, testbed.h:669). */
typedef struct insta_config {
    /* multiresolution hash grid ("encoding": HashGrid) */
    uint32_t n_levels;               /* L */
    uint32_t n_features_per_level;   /* F in {1, 2, 4, 8} */
    uint32_t log2_hashmap_size;      /* T = 2^log2_hashmap_size entries per level (cap) */
    uint32_t base_resolution;        /* N_min */
    float    per_level_scale;        /* b */
    /* bias-free ReLU MLPs of width n_neurons ("network", "rgb_network") */
    uint32_t n_neurons;              /* 64 */
    uint32_t density_hidden_layers;  /* >= 1 */
    uint32_t density_out;            /* 16: rgb input = density_out + SH coefficients */
    uint32_t rgb_hidden_layers;      /* >= 1 */
    uint32_t sh_degree;              /* 4 -> 16 coefficients */
    uint32_t n_extra_dims;           /* appended to the density input: 0, or 16
                                        expression values per frame on the
                                        device paths (exp_params of the
                                        reference's configs/nerf/main.json) */
    float    density_logit_bias;     /* sigma = exp(logit + bias); 0 = reference */
    /* occupancy grid and ray march (cone angle 0, one cascade) */
    uint32_t grid_size;              /* G: occupancy cells per axis over [0,1]^3 */
    float    dt;                     /* fixed step (reference: sqrt(3)/1024) */
    uint32_t max_steps;              /* cap on composited samples per ray */
    float    min_transmittance;      /* early termination (reference 0.01) */
    uint32_t disable_interp;         /* 1: never blend neighbour frames */
    uint32_t precision;              /* INSTA_PRECISION_* for tables/MLP */
    /* capacities */
    uint32_t max_frames;
    uint32_t max_triangles;          /* <= INSTA_MAX_TRIANGLES */
    uint32_t max_vertices;
    /* nearest-triangle BVH: 4^k count-balanced leaves of bvh_leaf_min ..
     * 4 * bvh_leaf_min - 1 triangles (the reference's leaves hold 2,
     * tiny_mesh.cu:28-32); a layout choice, the query result does not
     * depend on it.  In [1, 64]; default 4 */
    uint32_t bvh_leaf_min;
    /* nearest-triangle candidate lists: insta_build_occupancy also splits
     * each occupied cell of the frame's mesh occupancy grid into
     * cell_split^3 sub-cells and keeps, per sub-cell, every triangle that can
     * be the nearest one to a point inside it (its box within the sub-cell's
     * largest possible nearest distance), sorted by box distance.  Queries in
     * such a sub-cell scan the list; all others traverse the BVH.  The result
     * is the same either way.  0 = BVH only, else 1, 2 or 4 (default 4).  The
     * lists cost set-up per posed frame (512^2 head: 2.3 ms at 4, 1.4 at 2,
     * against 0.24 for the occupancy grid alone) and repay it when a frame is
     * rendered about three times or more; one render per pose is fastest at
     * 0, a served request's frame at 2 */
    uint32_t cell_split;
    uint32_t cell_list_cells;        /* occupied cells per frame that get lists (cap) */
    uint32_t cell_list_entries;      /* list entries per frame (cap; sub-cells past it use the BVH) */
} insta_config;

#define INSTA_MAX_TRIANGLES 16384u
#define INSTA_MAX_LEVELS 32u

/* Pinhole camera, instant-ngp convention: `matrix` is 3x4 row-major with
 * columns right, down, forward, position (common_device.cuh:202-222). */
typedef struct insta_camera {
    float matrix[12];
    float focal_px[2];
    float screen_center[2];          /* (0.5, 0.5) = image centre */
    uint32_t snap_to_pixel_centers;  /* 1: pixel centres, else ld jitter by spp */
} insta_camera;

typedef struct insta_render_stats {
    uint64_t n_rays;           /* rays initialised (pixels) */
    uint64_t n_rays_alive0;    /* rays that entered the render box */
    uint64_t n_samples;        /* samples composited (the deformed-sample count) */
    uint64_t n_samples_emitted;/* samples marched, deformed and evaluated */
    uint32_t n_iterations;     /* march/deform/network/composite rounds */
    uint32_t n_deform_misses;  /* samples inside the unit cube whose nearest-triangle query
                                  found nothing and were left undeformed (the reference's
                                  `id == -1` path of warp, src/core.cu:482-485) */
} insta_render_stats;

/* Parameter layout: density MLP weights, colour MLP weights, hash tables.
 * Weights are row-major [out][in]; the order mirrors NerfNetwork::set_params
 * (instant-ngp/include/neural-graphics-primitives/nerf_network.h:492-525). */
typedef struct insta_layout {
    uint64_t n_params;
    uint64_t n_mlp_params;
    uint64_t n_grid_params;        /* sum over levels of level_size * F */
    uint32_t density_in;           /* L*F + n_extra_dims */
    uint32_t rgb_in;               /* density_out + (sh_degree)^2 */
    uint32_t level_offset[INSTA_MAX_LEVELS];  /* entries before level l */
    uint32_t level_size[INSTA_MAX_LEVELS];    /* entries in level l */
    uint32_t level_resolution[INSTA_MAX_LEVELS];
    float    level_scale[INSTA_MAX_LEVELS];
    uint32_t level_dense[INSTA_MAX_LEVELS];   /* 1: dense index, 0: spatial hash */
} insta_layout;

const char* insta_last_error(void);
const char* insta_version(void);

/* Default configuration (the full-size workload shapes). */
void insta_default_config(insta_config* cfg);

/* Run settings of the reference's JSON configuration that are not network
 * shape: the optimizer (configs/nerf/main.json:21-32: Ema over Adam) and
 * INSTA's run keys (src/core.cu:199-221). */
typedef struct insta_run_config {
    float    learning_rate, beta1, beta2, epsilon;  /* Adam */
    float    l2_reg;              /* Adam's L2 regularisation (gradient += l2_reg * param) */
    float    ema_decay;           /* Ema wrapper decay; 0 = no Ema */
    uint32_t exp_params;          /* expression parameters = the density input's extra dims */
    uint32_t geo_params, use_eyes, optimize_latent_code;
    uint32_t max_training_steps;  /* "max_steps": training iterations of a run */
    float    depth_weight;        /* depth supervision lambda */
    float    seg_weight;          /* segmentation supervision lambda */
    uint32_t max_cached_bvh, max_images_gpu;
} insta_run_config;

void insta_default_run_config(insta_run_config* run);

/* Testbed::load_network_config / reload_network_from_file
 * (instant-ngp/src/testbed.cu:77-152) with the readers of reset_network
 * (testbed.cu:2181-2300), NerfNetwork (nerf_network.h:135-152) and
 * rta::Core::load_training_data (src/core.cu:199-221): reads a JSON file
 * (comments allowed), merges its "parent" chain (the parent, relative to the
 * child's directory, merge-patched with the child, RFC 7386), and fills cfg
 * (insta_default_config for absent keys; encoding, network, rgb_network,
 * dir_encoding, exp_params -> n_extra_dims, disable_interp; per_level_scale
 * from the finest resolution 2048 when absent) and run (nullable).  Returns
 * INSTA_ERR_UNSUPPORTED for an encoding, network, activation, direction
 * encoding, loss or optimizer this library does not implement.  Host only. */
insta_status insta_load_config_json(const char* path, insta_config* cfg, insta_run_config* run);

/* Parameter/table layout for a configuration (host-only, no device). */
insta_status insta_query_layout(const insta_config* cfg, insta_layout* out);

typedef struct insta_ctx insta_ctx;

/* Testbed::reset_network + rta::Core construction
 * (instant-ngp/src/testbed.cu:2181-2412, src/core.cu:199-221): allocates the
 * context on the current HIP device with capacities from `cfg`. */
insta_status insta_create(const insta_config* cfg, insta_ctx** out);
void insta_destroy(insta_ctx* ctx);

/* Canonical mesh + region masks + edge adjacency
 * (src/core.cu:283-304, src/tiny_mesh.cu:134-164, 316-371).  Host arrays:
 * verts [nv*3] float (already offset into the unit cube), faces [nf*3] int32,
 * mask_ids [nf] uint32 (FlameMasks values).  Also fixes the render box:
 * canonical vertex AABB inflated by 0.15 * |diag|, clipped to [0,1]^3
 * (tiny_mesh.cu:157-162, core.cu:289, testbed_nerf.cu:2675-2682). */
insta_status insta_set_canonical_mesh(insta_ctx* ctx, const float* verts, uint32_t nv,
                                      const int32_t* faces, uint32_t nf,
                                      const uint32_t* mask_ids, void* stream);

/* Posed frames: per-frame triangles, device BVH build and FP64-built
 * per-triangle frames, replacing the per-frame TinyMesh + CPU
 * TriangleBvh::build (src/core.cu:309-331, src/tiny_mesh.cu:134-178,
 * instant-ngp/src/triangle_bvh.cu:548-631).  `verts_dev` holds n_frames
 * consecutive [nv*3] float vertex arrays on the device (same topology as the
 * canonical mesh); frames first_frame .. first_frame+n_frames-1 are rebuilt in
 * one launch. */
insta_status insta_set_frames(insta_ctx* ctx, uint32_t first_frame, uint32_t n_frames,
                              const float* verts_dev, void* stream);

/* Expression blend on the device: frame vertices = canonical + sum_k
 * coeffs[k] * bases[k] (bases_dev [K][nv][3]), then the frame's BVH/frames are
 * rebuilt as insta_set_frames does.  Serving path for per-request expressions. */
insta_status insta_set_frame_expression(insta_ctx* ctx, uint32_t frame, const float* bases_dev,
                                        uint32_t n_bases, const float* coeffs_dev, void* stream);

/* Per-frame extra dims of the density input (n_extra_dims values per frame,
 * values_dev [n_frames][n_extra_dims] on the device), the expression
 * conditioning of Core::load_meshes (src/core.cu:251-274: the first
 * exp_params FLAME values of each frame).  A sample keeps its frame's values
 * when its nearest triangle is in the lips region (mask 7) or it is not
 * deformed, and takes 1.0 everywhere else (warp, src/core.cu:514-519); the
 * density-grid update feeds 1.0 throughout (extract_expressions_random,
 * nerf_network.h:107-127).  Rendering a frame uses that frame's values
 * (Core::render_deformed sets extra_dim_idx_for_inference to the frame,
 * src/core.cu:171).  Frames start at zero; insta_set_frame_expression also
 * sets them, to its first n_extra_dims coefficients (zero-padded). */
insta_status insta_set_frame_extra_dims(insta_ctx* ctx, uint32_t first_frame, uint32_t n_frames,
                                        const float* values_dev, void* stream);

/* Nearest triangle of each point on a frame's mesh (frame = -1: canonical),
 * the query of `closest_triangle`
 * (instant-ngp/include/neural-graphics-primitives/bvh.cuh:48-96).  points_dev
 * [n*3]; ids_dev [n] receives the original face index (-1 if none);
 * dist_dev [n] (nullable) receives the Euclidean distance. */
insta_status insta_closest_triangles(insta_ctx* ctx, int32_t frame, const float* points_dev,
                                     uint32_t n, int32_t* ids_dev, float* dist_dev, void* stream);

/* Candidate lists of a frame (see insta_config.cell_split), as left by its
 * last insta_build_occupancy: occupied cells, entries used, whether the lists
 * answer queries (0 when cell_split is 0, before the first build, after the
 * frame's mesh changed, or when the occupied cells exceeded cell_list_cells),
 * and sub-cells that fell back to the BVH because cell_list_entries ran out.
 * Waits for the stream.  No counterpart in the reference (its queries always
 * walk the BVH, bvh.cuh:48-96). */
typedef struct insta_cell_list_info {
    uint32_t occupied_cells;
    uint32_t entries;
    uint32_t valid;
    uint32_t subcells_without_list;
} insta_cell_list_info;
insta_status insta_get_cell_list_info(insta_ctx* ctx, int32_t frame, insta_cell_list_info* out, void* stream);

/* Export a frame's BVH (frame = -1: canonical) for inspection: nodes_dev
 * [n_nodes*8] floats (min xyz, begin, max xyz, end; begin/end are int32 bit
 * patterns, -1 for inner nodes), order_dev [nf] = original face index of
 * each BVH-order triangle.  Pass NULL pointers to query n_nodes only. */
insta_status insta_export_bvh(insta_ctx* ctx, int32_t frame, float* nodes_dev, uint32_t* order_dev,
                              uint32_t* n_nodes, void* stream);

/* Posed -> canonical deformation of NerfCoordinate samples in place: the
 * `warp` kernel behind Core::surface_closest_point and Core::render_deformed
 * (src/core.cu:442-526, 581-637).  frame = -1 takes each sample's frame from
 * its holder field; otherwise all samples use `frame`.  stride_floats =
 * 11 + n_extra_dims. */
insta_status insta_deform_samples(insta_ctx* ctx, float* coords_dev, uint32_t n,
                                  uint32_t stride_floats, int32_t frame, void* stream);

/* Occupancy bitfield from mesh distance: a cell (Morton order, G^3 cells over
 * [0,1]^3) is occupied when its centre is closer than `threshold` to the mesh
 * and the nearest triangle's mask id is not 0
 * (instant-ngp/src/testbed_nerf.cu:364-389, 585-609).  frame = -1 uses the
 * canonical mesh and fills the canonical grid slot, frame >= 0 the frame's
 * own slot.  bitfield_dev (nullable) receives a copy of the G^3/8 bytes. */
insta_status insta_build_occupancy(insta_ctx* ctx, int32_t frame, float threshold,
                                   uint8_t* bitfield_dev, void* stream);

/* Upload an occupancy bitfield (G^3/8 bytes, Morton order) into a slot. */
insta_status insta_set_occupancy(insta_ctx* ctx, int32_t slot, const uint8_t* bitfield_dev,
                                 void* stream);

/* Copy a slot's occupancy bitfield out (the density_grid_bitfield the march
 * reads, testbed_nerf.cu:2863-2878). */
insta_status insta_get_occupancy(insta_ctx* ctx, int32_t slot, uint8_t* bitfield_dev, void* stream);

/* NerfNetwork::set_params (nerf_network.h:492-525): params_dev holds
 * layout.n_params FP32 values in the declared order; the context keeps FP32
 * master copies and, for reduced precision, converted tables/weights. */
insta_status insta_set_params(insta_ctx* ctx, const float* params_dev, uint64_t n_params,
                              void* stream);

/* Hash-grid encoding forward (tiny-cuda-nn GridEncoding, absent from the
 * reference tree; called from nerf_network.h:168-174): positions_dev [n*3]
 * in [0,1]^3, features_dev [n * L*F] row-major (sample-major). */
insta_status insta_encode(insta_ctx* ctx, const float* positions_dev, uint32_t n,
                          float* features_dev, void* stream);

/* Hash-grid backward: grad_tables_dev[layout.n_grid_params] +=
 * sum over samples of dL/dfeature * trilinear weight (scatter-add).
 * dL_dfeatures_dev [n * L*F] row-major. */
insta_status insta_encode_backward(insta_ctx* ctx, const float* positions_dev, uint32_t n,
                                   const float* dL_dfeatures_dev, float* grad_tables_dev,
                                   void* stream);

/* NerfNetwork::inference_mixed_precision_impl (nerf_network.h:159-232) on
 * canonical NerfCoordinate samples: out_dev [n*4] = rgb logits (3) and the
 * density logit (density-MLP output 0). */
insta_status insta_network_inference(insta_ctx* ctx, const float* coords_dev, uint32_t n,
                                     uint32_t stride_floats, float* out_dev, void* stream);

/* ---- network plugin (NerfNetwork : tcnn::Network, nerf_network.h:131-652) --
 * The caller-driven forward/backward of the reference's network on
 * NerfCoordinate samples (stride_floats = 11 + n_extra_dims; the density
 * input's extra dims are each sample's last n_extra_dims floats).  They run
 * the training step's kernels: FP32 hash grid and FP32 MFMA MLPs whatever the
 * context's inference precision.  Shapes as insta_train_step supports
 * (F = 2, L = 8 or 16, width 64, 1 density / 2 colour hidden layers). */

/* forward_impl (nerf_network.h:238-294): out_dev [n*4] = rgb logits and the
 * density logit; keeps the activations the backward needs in the context
 * (one stored forward per context; insta_train_step overwrites it). */
insta_status insta_network_forward(insta_ctx* ctx, const float* coords_dev, uint32_t n, uint32_t stride_floats,
                                   float* out_dev, void* stream);
/* backward_impl (nerf_network.h:296-383) of the stored forward (same
 * coordinates and n): dL_doutput_dev [n*4] (rgb logits, density logit) ->
 * grad_params_dev [layout.n_params] in the insta_set_params order,
 * overwritten (accumulate = 0, EGradientMode::Overwrite) or added to
 * (accumulate = 1); dL_dinput_dev (nullable) [n * stride_floats] receives the
 * position rows 0..2 and the direction rows 8..10 (the input the
 * direction encoding reads, in [0,1]); other rows are left as they are. */
insta_status insta_network_backward(insta_ctx* ctx, const float* coords_dev, uint32_t n, uint32_t stride_floats,
                                    const float* dL_doutput_dev, float* grad_params_dev, int accumulate,
                                    float* dL_dinput_dev, void* stream);
/* density (nerf_network.h:385-414): positions_dev [n*3] -> out_dev [n*16],
 * the density MLP's 16 outputs (output 0 the density logit), extra dims 1.0
 * (extract_expressions_random, nerf_network.h:107-127). */
insta_status insta_network_density(insta_ctx* ctx, const float* positions_dev, uint32_t n, float* out_dev,
                                   void* stream);

/* The parameters renders and insta_network_inference use (n_params floats,
 * insta_get_params' order): the training weights, or their debiased average
 * while training runs with ema_decay > 0.  No counterpart beyond the
 * reference trainer's inference parameters (tcnn Trainer::params_inference). */
insta_status insta_get_inference_params(insta_ctx* ctx, float* params_dev, void* stream);

/* Workspace reservation.  insta_render, insta_render_frame, insta_train_step,
 * insta_train_loop and the network plugin entry points never allocate: their
 * buffers are sized here, ahead of the calls (the reference sizes its
 * scratch arenas per call, testbed_nerf.cu:2246-2271, 3176-3202).  Each field
 * grows its workspace to at least that size (0 leaves it as it is; nothing
 * shrinks).  A call that needs more than the reservation fails with
 * INSTA_ERR_STATE and names this entry point. */
typedef struct insta_reserve_config {
    uint32_t render_pixels;   /* width * height of the largest render */
    uint32_t train_rays;      /* rays per training step (1 << 18 covers the adaptive count) */
    uint32_t train_samples;   /* marched samples per step (insta_train_config.max_samples) */
    uint32_t train_batch;     /* composited batch (target_samples); also the network plugin's batch */
} insta_reserve_config;
insta_status insta_reserve(insta_ctx* ctx, const insta_reserve_config* r, void* stream);

/* Testbed::render_frame -> render_nerf -> NerfTracer::{init_rays_from_camera,
 * trace} with the render_deformed hook (testbed_nerf.cu:2030-2244,
 * 2320-2461; core.cu:346-348): renders `frame` through occupancy slot
 * `occupancy` (-1 canonical) at sample index `spp`.  rgba_dev [w*h*4]
 * (composited RGBA, 0 where the ray collected alpha <= 0.001),
 * depth_dev [w*h] (nullable; 1e10 where alpha <= 0.2), steps_dev [w*h]
 * (nullable; samples composited per pixel). */
insta_status insta_render(insta_ctx* ctx, const insta_camera* cam, uint32_t width, uint32_t height,
                          int32_t frame, int32_t occupancy, uint32_t spp, float* rgba_dev,
                          float* depth_dev, uint32_t* steps_dev, void* stream);

/* insta_render without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops with a fixed memory plan, harnesses that forbid hipMalloc).
 * Same arguments and result as insta_render; the round counters and the ray
 * and sample buffers live in one fixed device array of the library (about
 * 73 MiB per GPU, held from library load to process exit) instead of the
 * context's grown workspace.  At most 65536 pixels (256 x 256) per call
 * (INSTA_ERR_INVALID beyond); a march round holds at most 2^20 samples, so
 * larger frames take further rounds.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU race.  Calls no hipMalloc/hipFree. */
insta_status insta_render_noalloc(insta_ctx* ctx, const insta_camera* cam, uint32_t width, uint32_t height,
                                  int32_t frame, int32_t occupancy, uint32_t spp, float* rgba_dev,
                                  float* depth_dev, uint32_t* steps_dev, void* stream);

/* Counters of the last insta_render on this context. */
insta_status insta_get_render_stats(insta_ctx* ctx, insta_render_stats* out);

/* Samples the AoS deformation entry points (insta_deform_samples,
 * insta_deform_positions) left undeformed because their nearest-triangle
 * query found nothing (src/core.cu:482-485, 554-557), summed since the
 * context was created or last reset (reset != 0 zeroes it after reading).
 * Stream-ordered read (synchronises `stream`). */
insta_status insta_get_deform_misses(insta_ctx* ctx, uint64_t* misses, int reset, void* stream);

/* ---- render buffer (Testbed::render_frame's CudaRenderBuffer) -------- */

enum {
    INSTA_TONEMAP_IDENTITY = 0,  /* ETonemapCurve (render_buffer.h), the reference default */
    INSTA_TONEMAP_ACES = 1,
    INSTA_TONEMAP_HABLE = 2,
    INSTA_TONEMAP_REINHARD = 3
};

typedef struct insta_frame_config {
    float    exposure;        /* m_exposure (testbed.h:489): colour * 2^exposure in linear space */
    float    background[4];   /* m_background_color, sRGB + weight (testbed.h:817: 1,1,1,1) */
    uint32_t tonemap_curve;   /* INSTA_TONEMAP_* */
    uint32_t output_srgb;     /* render_frame(to_srgb): 1 sRGB output, 0 linear */
    uint32_t linear_colors;   /* Nerf::Training::linear_colors: 0 (default) -> shade converts the
                                 network's sRGB colour to linear before accumulating */
} insta_frame_config;

/* exposure 0, white background, identity curve, sRGB output, sRGB network colours */
void insta_default_frame_config(insta_frame_config* fc);

/* The render-buffer stages of Testbed::render_frame
 * (instant-ngp/src/testbed.cu:2613-2775) for one sample index `spp`, on a
 * composited RGBA image as insta_render writes it (rgba_dev [n*4], 0 where
 * alpha <= 0.001, depth_dev [n] nullable): shade_kernel_nerf over a cleared
 * frame buffer (sRGB -> linear unless linear_colors; depth buffer = depth
 * where alpha > 0.2, else 0; testbed_nerf.cu:1794-1827), accumulate_kernel
 * into accum_dev [n*4] (running mean over spp = 0, 1, 2, ...; spp 0 starts
 * afresh; render_buffer.cu:236-266, 560-578) and tonemap_kernel (background
 * blend in linear space, exposure, curve, output colour space;
 * render_buffer.cu:493-522) into out_dev [n*4].  As the reference's
 * tonemap_kernel does, the tonemapped colour is also written back into the
 * accumulation buffer, which the next sample index averages with.
 * depth_out_dev [n] (nullable) receives the depth buffer. */
insta_status insta_frame_buffer(const float* rgba_dev, const float* depth_dev, uint32_t n, uint32_t spp,
                                const insta_frame_config* fc, float* accum_dev, float* out_dev,
                                float* depth_out_dev, void* stream);

/* Testbed::render_frame: insta_render at sample index `spp` followed by
 * insta_frame_buffer (out_dev receives the tonemapped frame; accum_dev the
 * caller's accumulation buffer across sample indices). */
insta_status insta_render_frame(insta_ctx* ctx, const insta_camera* cam, uint32_t width, uint32_t height,
                                int32_t frame, int32_t occupancy, uint32_t spp, const insta_frame_config* fc,
                                float* accum_dev, float* out_dev, float* depth_out_dev, void* stream);

/* ---- training step --------------------------------------------------- */

/* Training set: n_images RGBA float images, rgba_dev [n][height][width][4]
 * (colour in [0,1]; alpha 1 = opaque, else composited over the random
 * background), image i seen by cameras[i] (host array) through posed frame
 * frames[i] (host array; that frame's mesh must be built, and its occupancy
 * grid too before a step whose rays march their frames' grids).  The context keeps its own copy of the images.  Stands
 * for the NeRF dataset load + Testbed::load_training_data
 * (instant-ngp/src/nerf_loader.cu:405-540, testbed.cu:2951-2980). */
insta_status insta_set_training_set(insta_ctx* ctx, uint32_t n_images, uint32_t width, uint32_t height,
                                    const float* rgba_dev, const insta_camera* cameras,
                                    const int32_t* frames, void* stream);

/* Segmentation masks and depth images of the training set last given to
 * insta_set_training_set (which drops any earlier ones), either NULL:
 * seg_dev [n][height][width] uint8 mask ids (the first channel of the
 * reference's RGBA mask image), depth_dev [n][height][width] float depth along
 * the camera axis in scene units (the reference's depth texel times its depth
 * scale), 0 = no depth at that pixel.  The context keeps copies.  In the
 * loss (compute_loss_kernel_train_nerf, testbed_nerf.cu:1476-1497, 1555-1586):
 * rays whose mask id is 120, 130 or 140 have their colour gradient scaled by
 * seg_weight; with masks set, ids <= 90 or > 250 and texels of alpha < 0.001
 * get no depth term; otherwise, where the depth texel is positive, the L1 loss
 * depth_weight * |D - |d| z| joins the density gradient, D the composited
 * distance |x - o| of the samples, |d| the unnormalised ray direction's length,
 * z the depth texel.  Stands for the seg_mask_path / depth_path loads of
 * NerfDataset (instant-ngp/src/nerf_loader.cu:773-800, 992-1012). */
insta_status insta_set_training_supervision(insta_ctx* ctx, const uint8_t* seg_dev, const float* depth_dev,
                                            void* stream);

typedef struct insta_train_config {
    uint32_t n_rays;          /* rays this step; 0 = adaptive count (aims at target_samples) */
    uint32_t target_samples;  /* training batch: composited samples compacted per step (cap) */
    uint32_t max_samples;     /* cap on samples marched per step */
    float    learning_rate;   /* Adam */
    float    beta1, beta2, epsilon;
    float    loss_scale;      /* gradients are computed scaled by this, then unscaled */
    uint32_t grid_grad_fp16;  /* 1: hash-grid gradients accumulated in FP16 (packed atomics): per
                                 chunk of 8192 batch entries in LDS for levels of at most 16384
                                 entries (the chunks then summed in FP32), per XCD copy for larger
                                 levels; 0: FP32 throughout */
    uint32_t apply_update;    /* 0: gradients only, parameters unchanged */
    uint32_t seed;            /* PCG32 seed of the ray/background stream */
    int32_t  occupancy_slot;  /* grid the rays march: INSTA_OCCUPANCY_PER_FRAME (each image's
                                 frame's own grid) or one slot for all rays (-1 canonical,
                                 f frame f), e.g. the density-grid EMA's */
    float    l2_reg;          /* Adam's L2 term: l2_reg * w joins every parameter's gradient in
                                 the moment updates (configs/nerf/main.json "l2_reg" 1e-6) */
    float    ema_decay;       /* 0: off.  Otherwise the Ema optimizer around Adam
                                 (main.json "Ema", decay 0.95): after each update the
                                 average a = decay * a + (1 - decay) * w is kept, and renders,
                                 insta_network_inference and insta_get_inference_params use
                                 a / (1 - decay^t) (t updates since it started); training
                                 itself keeps using w */
    float    seg_weight;      /* colour-gradient weight of rays whose mask id is 120, 130 or 140
                                 (seg_mask_supervision_lambda, testbed.h:623, default 1;
                                 main.json "seg_weight" 40); used with insta_set_training_supervision */
    float    depth_weight;    /* weight of the L1 depth term (depth_supervision_lambda,
                                 testbed.h:622, default 1.5; main.json "depth_weight" 1.25) */
} insta_train_config;
#define INSTA_OCCUPANCY_PER_FRAME (-2)

typedef struct insta_train_stats {
    uint32_t n_rays;          /* rays generated */
    uint32_t n_rays_used;     /* rays that collected samples (and fit the cap) */
    uint32_t n_samples;       /* samples marched */
    uint32_t n_samples_used;  /* samples composited before transmittance < 1e-4 */
    float    loss;            /* mean Huber loss per ray (unscaled) */
    uint32_t step;            /* training steps taken (the reference's m_training_step, which
                                 schedules the density-grid updates of insta_train_loop) */
    uint32_t adam_step;       /* Adam's step count (bias correction); restarts at 0 with
                                 fresh moments, e.g. a snapshot loaded without them */
    uint32_t n_samples_evaluated; /* of n_samples, those deformed and evaluated: the
                                 step marches every ray in full but evaluates its
                                 samples in windows, and only while the ray's
                                 transmittance is still >= 1e-4 (the loss reads
                                 no sample past that) */
} insta_train_stats;

/* Defaults: adaptive rays aiming at a 2^18-sample batch, 2^22 marched samples
 * at most, Adam (0.9, 0.99, 1e-15) at learning rate 1e-2, FP32 grid
 * gradients, loss scale 1, seed 1337, per-frame occupancy grids. */
void insta_default_train_config(insta_train_config* tc);


/* One training step, Testbed::train_nerf_step (testbed_nerf.cu:3176-3400):
 * rays from random training pixels (generate_training_samples_nerf,
 * 1105-1285) marched through their frame's occupancy grid, deformed to
 * canonical space (the warp hook, core.cu:442-526), network inference, loss
 * and dL/doutput (compute_loss_kernel_train_nerf, 1305-1643: Huber/5 at 0.1,
 * product transmittance, random background) with the composited samples
 * compacted into the training batch, FP32 forward and hand-written backward
 * through both MLPs and the hash grid on that batch (nerf_network.h:296-383),
 * Adam.  Reads one block of counters back to the host (the adaptive ray
 * count, stats). */
insta_status insta_train_step(insta_ctx* ctx, const insta_train_config* tc, insta_train_stats* stats,
                              void* stream);

/* dL/dparams of the last step (loss scale removed), layout.n_params floats. */
insta_status insta_get_gradients(insta_ctx* ctx, float* grads_dev, void* stream);

/* Current FP32 parameters (layout.n_params), e.g. after training steps. */
insta_status insta_get_params(insta_ctx* ctx, float* params_dev, void* stream);

/* Zero the Adam moments and Adam's step count (stream-ordered); the training
 * step count is kept. */
insta_status insta_reset_optimizer(insta_ctx* ctx, void* stream);

/* warp_random_samples (src/core.cu:528-579): positions_dev [n][3] mapped in
 * place to canonical space, each through a random posed frame: position i
 * advances the PCG32 stream seeded by `seed` (then advanced by rng_offset
 * draws) by i * (n_frames - 1) draws and takes frame floor(u * (n_frames -
 * 1)), so frame n_frames - 1 is never used (the reference's off-by-one,
 * kept).  Frames 0 .. n_frames - 2 must be built. */
insta_status insta_deform_positions(insta_ctx* ctx, float* positions_dev, uint32_t n, uint32_t n_frames,
                                    uint32_t seed, uint64_t rng_offset, void* stream);

typedef struct insta_density_grid_config {
    float    decay;            /* EMA decay (reference 0.95) */
    uint32_t n_uniform;        /* cells sampled at any density */
    uint32_t n_nonuniform;     /* cells sampled among those above 0.01 */
    uint32_t seed;             /* PCG32 seed of the cell/position and frame streams */
} insta_density_grid_config;

/* One density-grid update of training (Testbed::update_density_grid_nerf,
 * instant-ngp/src/testbed_nerf.cu:2774-2861): sample cells, warp the samples
 * to canonical space through random frames 0 .. n_frames - 2, splat the
 * density MLP's optical thickness (max per cell), EMA into the context's
 * float grid, then threshold at min(0.01, mean) into occupancy slot `slot`
 * (-1: the canonical slot, f: frame f's slot) for the march.  grid_dev
 * (nullable) receives the G^3 floats (Morton order); mean (nullable, host)
 * the grid mean (synchronises).  With a training set loaded, the first update
 * (and any after the number of training images changed) first marks the
 * cells no training camera sees as untrained, -1: never sampled, never
 * occupied (mark_untrained_density_grid, testbed_nerf.cu:391-431, 2795-2806). */
insta_status insta_update_density_grid(insta_ctx* ctx, int32_t slot, uint32_t n_frames,
                                       const insta_density_grid_config* cfg, float* grid_dev, float* mean,
                                       void* stream);

/* n_steps of the reference's training loop (Testbed::train +
 * training_prep_nerf, instant-ngp/src/testbed.cu:2480-2492,
 * testbed_nerf.cu:3428-3444): before optimizer step s the density grid is
 * updated when s % clamp(s / 16, 1, 16) == 0 -- G^3 uniform samples while
 * s < 256, then G^3/4 uniform + G^3/4 above the threshold -- with
 * dg->decay and dg->seed (its sample counts are ignored), through frames
 * 0 .. n_frames - 2 (warp_random_samples), into occupancy slot density_slot,
 * which the step's rays then march (tc->occupancy_slot is replaced by it).
 * last (nullable) receives the last step's statistics. */
insta_status insta_train_loop(insta_ctx* ctx, const insta_train_config* tc, uint32_t n_steps,
                              const insta_density_grid_config* dg, int32_t density_slot, uint32_t n_frames,
                              insta_train_stats* last, void* stream);
/* Restart the EMA (next update zeroes the grid and reseeds). */
insta_status insta_reset_density_grid(insta_ctx* ctx);

/* ---- serving -------------------------------------------------------- */

/* Request loop over several resident avatars on one device.  No counterpart
 * in the reference, which renders one frame at a time (Testbed::render_frame,
 * instant-ngp/src/testbed.cu:2613); each request runs the reference's render
 * path for a new expression: insta_set_frame_expression (device BVH),
 * insta_build_occupancy, insta_render.  n_slots requests are in flight at
 * once (a HIP stream and host thread each, overlapping on the device); the
 * FIFO queue head is admitted when a slot is idle and its render workspace
 * fits the HBM budget (one device arena, reserved at creation).  A request's
 * result depends only on its inputs, alone or in a mix. */
typedef struct insta_server insta_server;

typedef struct insta_server_config {
    uint32_t n_slots;              /* requests in flight at once */
    uint64_t hbm_budget_bytes;     /* device arena for the in-flight render workspaces */
    uint32_t samples_per_pixel;    /* sample buffer per pixel (more march rounds beyond it) */
    float    occupancy_threshold;  /* per-request occupancy grid: distance to the posed mesh */
    uint32_t use_graphs;           /* 1: per-request preparation replayed from captured hipGraphs */
} insta_server_config;

typedef struct insta_request {
    uint32_t avatar;               /* id from insta_server_add_avatar */
    const float* coeffs;           /* host [n_bases] expression coefficients, copied at submit */
    insta_camera camera;
    uint32_t width, height;
    float* rgba_dev;               /* caller-owned [w*h*4] output, written when the request completes */
    uint64_t user_tag;
} insta_request;

typedef struct insta_result {
    uint64_t ticket, user_tag;
    double t_submit, t_admit, t_done;  /* seconds on the server clock */
    uint64_t n_samples;            /* samples composited */
    int32_t status;                /* insta_status of the request */
    uint32_t slot;                 /* worker that ran it */
} insta_result;

/* 6 slots, 24 GiB budget, 64 samples per pixel, occupancy threshold 0.05, graphs on. */
void insta_default_server_config(insta_server_config* cfg);
insta_status insta_server_create(const insta_server_config* cfg, insta_server** out);
/* Register an avatar context (caller-owned, outliving the server; canonical
 * mesh and parameters set, max_frames >= n_slots: worker i uses frame slot i)
 * with its expression bases bases_dev [n_bases][nv][3].  Before any submit. */
insta_status insta_server_add_avatar(insta_server* srv, insta_ctx* ctx, const float* bases_dev,
                                     uint32_t n_bases, uint32_t* avatar_id);
/* Queue a request (never blocks on the device). */
insta_status insta_server_submit(insta_server* srv, const insta_request* req, uint64_t* ticket);
/* Take up to max_results finished requests; wait != 0 blocks until one is
 * finished or nothing is pending. */
insta_status insta_server_poll(insta_server* srv, insta_result* out, uint32_t max_results,
                               uint32_t* n_out, int wait);
/* Seconds since the server was created (the clock of insta_result). */
double insta_server_clock(insta_server* srv);
/* With use_graphs: requests whose captured chain (preparation + queued
 * render rounds) updated the worker's executable graph for that avatar in
 * place, and those that instantiated one (first request of a worker and
 * avatar, or a changed graph shape).  No counterpart in the reference. */
insta_status insta_server_graph_stats(insta_server* srv, uint64_t* updates, uint64_t* builds);
/* Finishes every queued request, then releases the server. */
void insta_server_destroy(insta_server* srv);

/* Snapshot of a trained avatar: Testbed::save_snapshot / load_snapshot
 * (instant-ngp/src/testbed.cu:2916-3012) for this path -- a msgpack map with
 * the network configuration ("encoding", "network", "rgb_network",
 * "dir_encoding", and "insta": every insta_config field) and "snapshot" with
 * the reference's keys (version, n_params, params_type "float",
 * params_binary in the insta_set_params order, density_grid_size,
 * density_grid_binary as FP16 in Morton order -- empty if the density grid
 * was never updated -- nerf.rgb.rays_per_batch, training_step, aabb,
 * render_aabb) and, with include_optimizer_state, the Adam moments.  Meshes
 * and the training set are not part of it (the reference reloads them from
 * the dataset).  Loading needs a context of the same network shape (see
 * insta_snapshot_config); the density grid's mean and bitfield are rebuilt
 * into occupancy slot occupancy_slot (-1 canonical, f posed frame f). */
insta_status insta_save_snapshot(insta_ctx* ctx, const char* path, int include_optimizer_state, void* stream);
insta_status insta_load_snapshot(insta_ctx* ctx, const char* path, int32_t occupancy_slot, void* stream);
/* The configuration stored in a snapshot (host only, no device work). */
insta_status insta_snapshot_config(const char* path, insta_config* cfg);

/* Per-kernel timing of insta_render, measured with HIP events recorded on the
 * render stream around each kernel class.  No counterpart in the reference
 * (whose only timers are EMA frame times, testbed.cu:1600, 2484-2512); it
 * exists so throughput can be divided among kernels.  Classes: */
enum {
    INSTA_KERNEL_INIT = 0,      /* ray init + jitter + empty-space skip */
    INSTA_KERNEL_MARCH = 1,     /* sample emission */
    INSTA_KERNEL_DEFORM = 2,    /* BVH nearest triangle + frame transform */
    INSTA_KERNEL_NETWORK = 3,   /* hash encoding + density/colour MLPs */
    INSTA_KERNEL_COMPOSITE = 4, /* front-to-back compositing + compaction */
    INSTA_KERNEL_FINALIZE = 5,  /* framebuffer write */
    INSTA_KERNEL_CLASSES = 8
};
typedef struct insta_profile {
    double ms[INSTA_KERNEL_CLASSES];         /* summed event time */
    uint64_t launches[INSTA_KERNEL_CLASSES];
    uint64_t items[INSTA_KERNEL_CLASSES];    /* rays or samples processed */
    /* with enable == 2: nearest-triangle work of the sample deformation
     * (render and training): queries, inner nodes expanded (4 child boxes
     * each), leaves visited, exact triangle distances evaluated (those whose
     * own box, and on the list path also whose plane, passed the bound) */
    uint64_t bvh_queries, bvh_inner, bvh_leaves, bvh_triangles;
    /* of bvh_queries, those answered from a candidate list, and the list
     * entries they scanned (their exact distances are in bvh_triangles) */
    uint64_t list_queries, list_entries;
} insta_profile;
/* enable != 0 starts (and zeroes) accumulation; 0 stops it.  enable == 2 also
 * counts traversal work, with a counting build of the deformation kernel
 * (slower: for a separate, untimed render). */
insta_status insta_set_profiling(insta_ctx* ctx, int enable);
insta_status insta_get_profile(insta_ctx* ctx, insta_profile* out);

#ifdef __cplusplus
}
#endif

#endif /* INSTA_CAPI_H */
