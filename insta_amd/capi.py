"""ctypes binding of the C ABI in include/insta_capi.h.

Device buffers are torch tensors on the current HIP device (PyTorch is only
used for device memory and streams here); every call passes raw pointers
through the C ABI, exactly as a C/C++ caller would.  Loading fails loudly if
the HIP library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libinsta_amd.so"

INSTA_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "HIP", 3: "OOM", 4: "STATE", 5: "UNSUPPORTED"}
PRECISION_FP32, PRECISION_FP16, PRECISION_BF16, PRECISION_FP32_SCALAR = 0, 1, 2, 3
PRECISION_NAMES = {0: "fp32", 1: "fp16", 2: "bf16", 3: "fp32_scalar"}
MAX_LEVELS = 32
MAX_TRIANGLES = 16384


class InstaConfig(C.Structure):
    _fields_ = [
        ("n_levels", C.c_uint32), ("n_features_per_level", C.c_uint32), ("log2_hashmap_size", C.c_uint32),
        ("base_resolution", C.c_uint32), ("per_level_scale", C.c_float),
        ("n_neurons", C.c_uint32), ("density_hidden_layers", C.c_uint32), ("density_out", C.c_uint32),
        ("rgb_hidden_layers", C.c_uint32), ("sh_degree", C.c_uint32), ("n_extra_dims", C.c_uint32),
        ("density_logit_bias", C.c_float),
        ("grid_size", C.c_uint32), ("dt", C.c_float), ("max_steps", C.c_uint32), ("min_transmittance", C.c_float),
        ("disable_interp", C.c_uint32), ("precision", C.c_uint32),
        ("max_frames", C.c_uint32), ("max_triangles", C.c_uint32), ("max_vertices", C.c_uint32),
        ("bvh_leaf_min", C.c_uint32), ("cell_split", C.c_uint32), ("cell_list_cells", C.c_uint32),
        ("cell_list_entries", C.c_uint32),
    ]


class InstaCellListInfo(C.Structure):
    _fields_ = [("occupied_cells", C.c_uint32), ("entries", C.c_uint32), ("valid", C.c_uint32),
                ("subcells_without_list", C.c_uint32)]


class InstaReserveConfig(C.Structure):
    _fields_ = [("render_pixels", C.c_uint32), ("train_rays", C.c_uint32), ("train_samples", C.c_uint32),
                ("train_batch", C.c_uint32)]


class InstaCamera(C.Structure):
    _fields_ = [("matrix", C.c_float * 12), ("focal_px", C.c_float * 2), ("screen_center", C.c_float * 2),
                ("snap_to_pixel_centers", C.c_uint32)]


class InstaRenderStats(C.Structure):
    _fields_ = [("n_rays", C.c_uint64), ("n_rays_alive0", C.c_uint64), ("n_samples", C.c_uint64),
                ("n_samples_emitted", C.c_uint64), ("n_iterations", C.c_uint32), ("n_deform_misses", C.c_uint32)]


TONEMAP_IDENTITY, TONEMAP_ACES, TONEMAP_HABLE, TONEMAP_REINHARD = 0, 1, 2, 3


class InstaRunConfig(C.Structure):
    _fields_ = [("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("l2_reg", C.c_float), ("ema_decay", C.c_float), ("exp_params", C.c_uint32), ("geo_params", C.c_uint32),
                ("use_eyes", C.c_uint32), ("optimize_latent_code", C.c_uint32), ("max_training_steps", C.c_uint32),
                ("depth_weight", C.c_float), ("seg_weight", C.c_float), ("max_cached_bvh", C.c_uint32),
                ("max_images_gpu", C.c_uint32)]


class InstaFrameConfig(C.Structure):
    _fields_ = [("exposure", C.c_float), ("background", C.c_float * 4), ("tonemap_curve", C.c_uint32),
                ("output_srgb", C.c_uint32), ("linear_colors", C.c_uint32)]


class InstaLayout(C.Structure):
    _fields_ = [
        ("n_params", C.c_uint64), ("n_mlp_params", C.c_uint64), ("n_grid_params", C.c_uint64),
        ("density_in", C.c_uint32), ("rgb_in", C.c_uint32),
        ("level_offset", C.c_uint32 * MAX_LEVELS), ("level_size", C.c_uint32 * MAX_LEVELS),
        ("level_resolution", C.c_uint32 * MAX_LEVELS), ("level_scale", C.c_float * MAX_LEVELS),
        ("level_dense", C.c_uint32 * MAX_LEVELS),
    ]

    def mlp_shapes(self, cfg: InstaConfig) -> list[tuple[int, int]]:
        """(out, in) of every weight matrix in parameter order."""
        w = cfg.n_neurons
        shapes = [(w, self.density_in)] + [(w, w)] * (cfg.density_hidden_layers - 1) + [(cfg.density_out, w)]
        shapes += [(w, self.rgb_in)] + [(w, w)] * (cfg.rgb_hidden_layers - 1) + [(3, w)]
        return shapes


KERNEL_CLASSES = ["init", "march", "deform", "network", "composite", "finalize"]


class InstaTrainConfig(C.Structure):
    _fields_ = [("n_rays", C.c_uint32), ("target_samples", C.c_uint32), ("max_samples", C.c_uint32),
                ("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("loss_scale", C.c_float), ("grid_grad_fp16", C.c_uint32), ("apply_update", C.c_uint32),
                ("seed", C.c_uint32), ("occupancy_slot", C.c_int32), ("l2_reg", C.c_float), ("ema_decay", C.c_float),
                ("seg_weight", C.c_float), ("depth_weight", C.c_float)]


class InstaTrainStats(C.Structure):
    _fields_ = [("n_rays", C.c_uint32), ("n_rays_used", C.c_uint32), ("n_samples", C.c_uint32),
                ("n_samples_used", C.c_uint32), ("loss", C.c_float), ("step", C.c_uint32),
                ("adam_step", C.c_uint32), ("n_samples_evaluated", C.c_uint32)]


class InstaDensityGridConfig(C.Structure):
    _fields_ = [("decay", C.c_float), ("n_uniform", C.c_uint32), ("n_nonuniform", C.c_uint32), ("seed", C.c_uint32)]


class InstaServerConfig(C.Structure):
    _fields_ = [("n_slots", C.c_uint32), ("hbm_budget_bytes", C.c_uint64), ("samples_per_pixel", C.c_uint32),
                ("occupancy_threshold", C.c_float), ("use_graphs", C.c_uint32)]


class InstaRequest(C.Structure):
    _fields_ = [("avatar", C.c_uint32), ("coeffs", C.c_void_p), ("camera", InstaCamera), ("width", C.c_uint32),
                ("height", C.c_uint32), ("rgba_dev", C.c_void_p), ("user_tag", C.c_uint64)]


class InstaResult(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("user_tag", C.c_uint64), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("n_samples", C.c_uint64), ("status", C.c_int32), ("slot", C.c_uint32)]


class InstaProfile(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("launches", C.c_uint64 * 8), ("items", C.c_uint64 * 8),
                ("bvh_queries", C.c_uint64), ("bvh_inner", C.c_uint64), ("bvh_leaves", C.c_uint64),
                ("bvh_triangles", C.c_uint64), ("list_queries", C.c_uint64), ("list_entries", C.c_uint64)]


# the C ABI: (name, restype, argtypes); tests check this list against the header
_VP = C.c_void_p
SIGNATURES = [
    ("insta_last_error", C.c_char_p, []),
    ("insta_version", C.c_char_p, []),
    ("insta_default_config", None, [C.POINTER(InstaConfig)]),
    ("insta_query_layout", C.c_int, [C.POINTER(InstaConfig), C.POINTER(InstaLayout)]),
    ("insta_default_run_config", None, [C.POINTER(InstaRunConfig)]),
    ("insta_load_config_json", C.c_int, [C.c_char_p, C.POINTER(InstaConfig), C.POINTER(InstaRunConfig)]),
    ("insta_create", C.c_int, [C.POINTER(InstaConfig), C.POINTER(_VP)]),
    ("insta_reserve", C.c_int, [_VP, C.POINTER(InstaReserveConfig), _VP]),
    ("insta_get_inference_params", C.c_int, [_VP, _VP, _VP]),
    ("insta_destroy", None, [_VP]),
    ("insta_set_canonical_mesh", C.c_int, [_VP, _VP, C.c_uint32, _VP, C.c_uint32, _VP, _VP]),
    ("insta_set_frames", C.c_int, [_VP, C.c_uint32, C.c_uint32, _VP, _VP]),
    ("insta_set_frame_expression", C.c_int, [_VP, C.c_uint32, _VP, C.c_uint32, _VP, _VP]),
    ("insta_set_frame_extra_dims", C.c_int, [_VP, C.c_uint32, C.c_uint32, _VP, _VP]),
    ("insta_closest_triangles", C.c_int, [_VP, C.c_int32, _VP, C.c_uint32, _VP, _VP, _VP]),
    ("insta_get_cell_list_info", C.c_int, [_VP, C.c_int32, C.POINTER(InstaCellListInfo), _VP]),
    ("insta_export_bvh", C.c_int, [_VP, C.c_int32, _VP, _VP, C.POINTER(C.c_uint32), _VP]),
    ("insta_deform_samples", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, C.c_int32, _VP]),
    ("insta_build_occupancy", C.c_int, [_VP, C.c_int32, C.c_float, _VP, _VP]),
    ("insta_set_occupancy", C.c_int, [_VP, C.c_int32, _VP, _VP]),
    ("insta_set_params", C.c_int, [_VP, _VP, C.c_uint64, _VP]),
    ("insta_encode", C.c_int, [_VP, _VP, C.c_uint32, _VP, _VP]),
    ("insta_encode_backward", C.c_int, [_VP, _VP, C.c_uint32, _VP, _VP, _VP]),
    ("insta_network_inference", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, _VP, _VP]),
    ("insta_network_forward", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, _VP, _VP]),
    ("insta_network_backward", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, _VP, _VP, C.c_int, _VP, _VP]),
    ("insta_network_density", C.c_int, [_VP, _VP, C.c_uint32, _VP, _VP]),
    ("insta_render", C.c_int, [_VP, C.POINTER(InstaCamera), C.c_uint32, C.c_uint32, C.c_int32, C.c_int32,
                               C.c_uint32, _VP, _VP, _VP, _VP]),
    ("insta_render_noalloc", C.c_int, [_VP, C.POINTER(InstaCamera), C.c_uint32, C.c_uint32, C.c_int32, C.c_int32,
                                       C.c_uint32, _VP, _VP, _VP, _VP]),
    ("insta_get_render_stats", C.c_int, [_VP, C.POINTER(InstaRenderStats)]),
    ("insta_get_deform_misses", C.c_int, [_VP, C.POINTER(C.c_uint64), C.c_int, _VP]),
    ("insta_default_frame_config", None, [C.POINTER(InstaFrameConfig)]),
    ("insta_frame_buffer", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, C.POINTER(InstaFrameConfig), _VP, _VP, _VP,
                                     _VP]),
    ("insta_render_frame", C.c_int, [_VP, C.POINTER(InstaCamera), C.c_uint32, C.c_uint32, C.c_int32, C.c_int32,
                                     C.c_uint32, C.POINTER(InstaFrameConfig), _VP, _VP, _VP, _VP]),
    ("insta_save_snapshot", C.c_int, [_VP, C.c_char_p, C.c_int, _VP]),
    ("insta_load_snapshot", C.c_int, [_VP, C.c_char_p, C.c_int32, _VP]),
    ("insta_snapshot_config", C.c_int, [C.c_char_p, C.POINTER(InstaConfig)]),
    ("insta_set_profiling", C.c_int, [_VP, C.c_int]),
    ("insta_get_profile", C.c_int, [_VP, C.POINTER(InstaProfile)]),
    ("insta_set_training_set", C.c_int, [_VP, C.c_uint32, C.c_uint32, C.c_uint32, _VP, _VP, _VP, _VP]),
    ("insta_set_training_supervision", C.c_int, [_VP, _VP, _VP, _VP]),
    ("insta_default_train_config", None, [C.POINTER(InstaTrainConfig)]),
    ("insta_train_step", C.c_int, [_VP, C.POINTER(InstaTrainConfig), C.POINTER(InstaTrainStats), _VP]),
    ("insta_train_loop", C.c_int, [_VP, C.POINTER(InstaTrainConfig), C.c_uint32, C.POINTER(InstaDensityGridConfig),
                                   C.c_int32, C.c_uint32, C.POINTER(InstaTrainStats), _VP]),
    ("insta_get_gradients", C.c_int, [_VP, _VP, _VP]),
    ("insta_get_params", C.c_int, [_VP, _VP, _VP]),
    ("insta_reset_optimizer", C.c_int, [_VP, _VP]),
    ("insta_get_occupancy", C.c_int, [_VP, C.c_int32, _VP, _VP]),
    ("insta_deform_positions", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, _VP]),
    ("insta_update_density_grid", C.c_int, [_VP, C.c_int32, C.c_uint32, C.POINTER(InstaDensityGridConfig), _VP,
                                            C.POINTER(C.c_float), _VP]),
    ("insta_reset_density_grid", C.c_int, [_VP]),
    ("insta_default_server_config", None, [C.POINTER(InstaServerConfig)]),
    ("insta_server_create", C.c_int, [C.POINTER(InstaServerConfig), C.POINTER(_VP)]),
    ("insta_server_add_avatar", C.c_int, [_VP, _VP, _VP, C.c_uint32, C.POINTER(C.c_uint32)]),
    ("insta_server_submit", C.c_int, [_VP, C.POINTER(InstaRequest), C.POINTER(C.c_uint64)]),
    ("insta_server_poll", C.c_int, [_VP, C.POINTER(InstaResult), C.c_uint32, C.POINTER(C.c_uint32), C.c_int]),
    ("insta_server_clock", C.c_double, [_VP]),
    ("insta_server_graph_stats", C.c_int, [_VP, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("insta_server_destroy", None, [_VP]),
]

_lib = None


def load_library(path: Path | str = LIB_PATH) -> C.CDLL:
    """Load the HIP library; raises if it is missing (build it first)."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path)
    if not p.exists():
        raise RuntimeError(f"{p} is missing: run `python -c 'import __graft_entry__ as g; g.build()'` first")
    lib = C.CDLL(str(p))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class InstaError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def check(status: int, where: str) -> None:
    if status != INSTA_OK:
        raise InstaError(status, where, load_library().insta_last_error().decode())


# ----------------------------------------------------------------------------
# workload configurations
# ----------------------------------------------------------------------------

def default_config() -> InstaConfig:
    cfg = InstaConfig()
    load_library().insta_default_config(C.byref(cfg))
    return cfg


def make_config(**kw) -> InstaConfig:
    """Full default config with overrides; fields are set verbatim."""
    cfg = InstaConfig()
    base = dict(n_levels=16, n_features_per_level=2, log2_hashmap_size=19, base_resolution=16,
                per_level_scale=math.exp(math.log(2048.0 / 16.0) / 15.0), n_neurons=64,
                density_hidden_layers=1, density_out=16, rgb_hidden_layers=2, sh_degree=4, n_extra_dims=0,
                density_logit_bias=1.0, grid_size=128, dt=1.73205080757 / 1024.0, max_steps=1024,
                min_transmittance=0.01, disable_interp=0, precision=PRECISION_FP32, max_frames=16,
                max_triangles=MAX_TRIANGLES, max_vertices=1 << 16, bvh_leaf_min=4, cell_split=4,
                cell_list_cells=1 << 16, cell_list_entries=1 << 25)
    base.update(kw)
    for k, v in base.items():
        setattr(cfg, k, v)
    return cfg


def small_config(**kw) -> InstaConfig:
    """Small workload: L=8 F=2 T=2^14 N_min=16 b=1.5, 64^3 grid, step 1/256,
    at most 256 steps, MLP 16->64->16 / 32->64->64->3."""
    base = dict(n_levels=8, log2_hashmap_size=14, per_level_scale=1.5, grid_size=64, dt=1.0 / 256.0,
                max_steps=256, max_frames=4, max_triangles=4096, max_vertices=4096)
    base.update(kw)
    return make_config(**base)


def full_config(**kw) -> InstaConfig:
    """Full-size workload: L=16 F=2 T=2^19, N_min=16 -> N_max=2048, 128^3
    grid, step sqrt(3)/1024, at most 1,024 steps, 16 frames."""
    return make_config(**kw)


def query_layout(cfg: InstaConfig) -> InstaLayout:
    lay = InstaLayout()
    check(load_library().insta_query_layout(C.byref(cfg), C.byref(lay)), "insta_query_layout")
    return lay


def make_camera(matrix: np.ndarray, focal: float, snap: bool = True, center=(0.5, 0.5)) -> InstaCamera:
    cam = InstaCamera()
    m = np.asarray(matrix, dtype=np.float32).reshape(3, 4)
    for i, v in enumerate(m.ravel()):
        cam.matrix[i] = float(v)
    cam.focal_px[0] = cam.focal_px[1] = float(focal)
    cam.screen_center[0], cam.screen_center[1] = center
    cam.snap_to_pixel_centers = 1 if snap else 0
    return cam


# ----------------------------------------------------------------------------
# context wrapper (device buffers are torch tensors)
# ----------------------------------------------------------------------------

def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


class Context:
    """One avatar on the current HIP device."""

    def __init__(self, cfg: InstaConfig):
        import torch
        self.torch = torch
        self.lib = load_library()
        self.cfg = cfg
        self.layout = query_layout(cfg)
        h = C.c_void_p()
        check(self.lib.insta_create(C.byref(cfg), C.byref(h)), "insta_create")
        self.h = h
        self.nv = self.nf = 0
        self.reserved = dict(render_pixels=0, train_rays=0, train_samples=0, train_batch=0)

    def reserve(self, **kw) -> None:
        """insta_reserve: grow the named workspaces (render_pixels,
        train_rays, train_samples, train_batch) to at least the given sizes."""
        r = InstaReserveConfig(**kw)
        check(self.lib.insta_reserve(self.h, C.byref(r), self.stream()), "insta_reserve")
        for k, v in kw.items():
            self.reserved[k] = max(self.reserved[k], v)

    def _reserve_for(self, **need) -> None:
        """Reserve ahead of a call whatever it needs beyond the reservation so
        far (the calls themselves never allocate)."""
        grow = {k: v for k, v in need.items() if v > self.reserved[k]}
        if grow:
            self.reserve(**grow)

    def close(self) -> None:
        if self.h:
            self.lib.insta_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream(self):
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def dev(self, a, dtype=None):
        t = self.torch.as_tensor(np.ascontiguousarray(a))
        if dtype is not None:
            t = t.to(dtype)
        return t.cuda()

    # --- mesh -------------------------------------------------------------
    def set_canonical_mesh(self, verts: np.ndarray, faces: np.ndarray, masks: np.ndarray) -> None:
        v = np.ascontiguousarray(verts, dtype=np.float32)
        f = np.ascontiguousarray(faces, dtype=np.int32)
        m = np.ascontiguousarray(masks, dtype=np.uint32)
        self.nv, self.nf = len(v), len(f)
        check(self.lib.insta_set_canonical_mesh(self.h, v.ctypes.data, len(v), f.ctypes.data, len(f),
                                                m.ctypes.data, self.stream()), "insta_set_canonical_mesh")

    def set_frames(self, first: int, frames_dev) -> None:
        n = frames_dev.shape[0] if frames_dev.dim() == 3 else 1
        check(self.lib.insta_set_frames(self.h, first, n, _ptr(frames_dev.contiguous()), self.stream()),
              "insta_set_frames")

    def set_frame_expression(self, frame: int, bases_dev, coeffs_dev) -> None:
        check(self.lib.insta_set_frame_expression(self.h, frame, _ptr(bases_dev), bases_dev.shape[0],
                                                  _ptr(coeffs_dev), self.stream()), "insta_set_frame_expression")

    def set_frame_extra_dims(self, first: int, values_dev) -> None:
        """values_dev [n_frames, n_extra_dims] float32 on the device."""
        v = values_dev.contiguous()
        n = v.shape[0] if v.dim() == 2 else 1
        check(self.lib.insta_set_frame_extra_dims(self.h, first, n, _ptr(v), self.stream()),
              "insta_set_frame_extra_dims")

    def closest(self, frame: int, pts_dev):
        n = pts_dev.shape[0]
        ids = self.torch.empty(n, dtype=self.torch.int32, device=pts_dev.device)
        dist = self.torch.empty(n, dtype=self.torch.float32, device=pts_dev.device)
        check(self.lib.insta_closest_triangles(self.h, frame, _ptr(pts_dev), n, _ptr(ids), _ptr(dist),
                                               self.stream()), "insta_closest_triangles")
        return ids, dist

    def cell_list_info(self, frame: int) -> dict:
        info = InstaCellListInfo()
        check(self.lib.insta_get_cell_list_info(self.h, frame, C.byref(info), self.stream()),
              "insta_get_cell_list_info")
        return {name: getattr(info, name) for name, _ in InstaCellListInfo._fields_}

    def export_bvh(self, frame: int):
        nn = C.c_uint32()
        check(self.lib.insta_export_bvh(self.h, frame, None, None, C.byref(nn), self.stream()), "insta_export_bvh")
        nodes = self.torch.empty(nn.value * 8, dtype=self.torch.float32, device="cuda")
        order = self.torch.empty(self.nf, dtype=self.torch.int32, device="cuda")
        check(self.lib.insta_export_bvh(self.h, frame, _ptr(nodes), _ptr(order), C.byref(nn), self.stream()),
              "insta_export_bvh")
        return nodes, order

    def deform_samples(self, coords_dev, frame: int = -1) -> None:
        n, stride = coords_dev.shape
        check(self.lib.insta_deform_samples(self.h, _ptr(coords_dev), n, stride, frame, self.stream()),
              "insta_deform_samples")

    def build_occupancy(self, frame: int, threshold: float):
        G = self.cfg.grid_size
        bits = self.torch.empty(G * G * G // 8, dtype=self.torch.uint8, device="cuda")
        check(self.lib.insta_build_occupancy(self.h, frame, threshold, _ptr(bits), self.stream()),
              "insta_build_occupancy")
        return bits

    def get_occupancy(self, slot: int):
        G = self.cfg.grid_size
        bits = self.torch.empty(G * G * G // 8, dtype=self.torch.uint8, device="cuda")
        check(self.lib.insta_get_occupancy(self.h, slot, _ptr(bits), self.stream()), "insta_get_occupancy")
        return bits

    def set_occupancy(self, slot: int, bits_dev) -> None:
        check(self.lib.insta_set_occupancy(self.h, slot, _ptr(bits_dev), self.stream()), "insta_set_occupancy")

    # --- network ----------------------------------------------------------
    def set_params(self, params_dev) -> None:
        check(self.lib.insta_set_params(self.h, _ptr(params_dev), params_dev.numel(), self.stream()),
              "insta_set_params")

    # --- snapshot -------------------------------------------------------------
    def save_snapshot(self, path, include_optimizer: bool = False) -> None:
        check(self.lib.insta_save_snapshot(self.h, str(path).encode(), 1 if include_optimizer else 0, self.stream()),
              "insta_save_snapshot")

    def load_snapshot(self, path, occupancy_slot: int = -1) -> None:
        check(self.lib.insta_load_snapshot(self.h, str(path).encode(), occupancy_slot, self.stream()),
              "insta_load_snapshot")

    def encode(self, pos_dev):
        n = pos_dev.shape[0]
        out = self.torch.empty((n, self.cfg.n_levels * self.cfg.n_features_per_level), dtype=self.torch.float32,
                               device=pos_dev.device)
        check(self.lib.insta_encode(self.h, _ptr(pos_dev), n, _ptr(out), self.stream()), "insta_encode")
        return out

    def encode_backward(self, pos_dev, dL_dev, grad_dev) -> None:
        check(self.lib.insta_encode_backward(self.h, _ptr(pos_dev), pos_dev.shape[0], _ptr(dL_dev), _ptr(grad_dev),
                                             self.stream()), "insta_encode_backward")

    def network_inference(self, coords_dev):
        n, stride = coords_dev.shape
        out = self.torch.empty((n, 4), dtype=self.torch.float32, device=coords_dev.device)
        check(self.lib.insta_network_inference(self.h, _ptr(coords_dev), n, stride, _ptr(out), self.stream()),
              "insta_network_inference")
        return out

    # --- network plugin (NerfNetwork forward / backward / density) ----------
    def network_forward(self, coords_dev):
        n, stride = coords_dev.shape
        self._reserve_for(train_samples=n, train_batch=n)
        out = self.torch.empty((n, 4), dtype=self.torch.float32, device=coords_dev.device)
        check(self.lib.insta_network_forward(self.h, _ptr(coords_dev), n, stride, _ptr(out), self.stream()),
              "insta_network_forward")
        return out

    def network_backward(self, coords_dev, dL_dout_dev, grads_dev=None, accumulate: bool = False,
                         dL_dinput_dev=None):
        """Parameter gradients of the stored forward (a new tensor unless
        grads_dev is given); dL_dinput_dev [n, stride] receives the position
        and direction rows when given."""
        n, stride = coords_dev.shape
        if grads_dev is None:
            grads_dev = self.torch.empty(self.layout.n_params, dtype=self.torch.float32, device=coords_dev.device)
        check(self.lib.insta_network_backward(self.h, _ptr(coords_dev), n, stride, _ptr(dL_dout_dev), _ptr(grads_dev),
                                              1 if accumulate else 0, _ptr(dL_dinput_dev), self.stream()),
              "insta_network_backward")
        return grads_dev

    def network_density(self, pos_dev):
        out = self.torch.empty((pos_dev.shape[0], 16), dtype=self.torch.float32, device=pos_dev.device)
        check(self.lib.insta_network_density(self.h, _ptr(pos_dev), pos_dev.shape[0], _ptr(out), self.stream()),
              "insta_network_density")
        return out

    # --- render -----------------------------------------------------------
    def render(self, cam: InstaCamera, width: int, height: int, frame: int, occupancy: int, spp: int = 0,
               out=None):
        torch = self.torch
        if out is None:
            out = (torch.empty((height * width, 4), dtype=torch.float32, device="cuda"),
                   torch.empty(height * width, dtype=torch.float32, device="cuda"),
                   torch.empty(height * width, dtype=torch.int32, device="cuda"))
        rgba, depth, steps = out
        self._reserve_for(render_pixels=width * height)
        check(self.lib.insta_render(self.h, C.byref(cam), width, height, frame, occupancy, spp, _ptr(rgba),
                                    _ptr(depth), _ptr(steps), self.stream()), "insta_render")
        return out

    def render_noalloc(self, cam: InstaCamera, width: int, height: int, frame: int, occupancy: int, spp: int = 0,
                       out=None):
        """insta_render_noalloc: the render on the library's fixed device
        array (no allocation in the call; at most 256 x 256 pixels)."""
        torch = self.torch
        if out is None:
            out = (torch.empty((height * width, 4), dtype=torch.float32, device="cuda"),
                   torch.empty(height * width, dtype=torch.float32, device="cuda"),
                   torch.empty(height * width, dtype=torch.int32, device="cuda"))
        rgba, depth, steps = out
        check(self.lib.insta_render_noalloc(self.h, C.byref(cam), width, height, frame, occupancy, spp, _ptr(rgba),
                                            _ptr(depth), _ptr(steps), self.stream()), "insta_render_noalloc")
        return out

    def render_frame(self, cam: InstaCamera, width: int, height: int, frame: int, occupancy: int, spp: int,
                     fc: "InstaFrameConfig", accum, out=None, depth=None):
        """Testbed::render_frame at sample index spp: render, shade,
        accumulate into `accum` [h*w, 4], tonemap into `out` (returned)."""
        torch = self.torch
        if out is None:
            out = torch.empty((height * width, 4), dtype=torch.float32, device="cuda")
        self._reserve_for(render_pixels=width * height)
        check(self.lib.insta_render_frame(self.h, C.byref(cam), width, height, frame, occupancy, spp, C.byref(fc),
                                          _ptr(accum), _ptr(out), _ptr(depth), self.stream()), "insta_render_frame")
        return out

    def deform_misses(self, reset: bool = False) -> int:
        v = C.c_uint64()
        check(self.lib.insta_get_deform_misses(self.h, C.byref(v), 1 if reset else 0, self.stream()),
              "insta_get_deform_misses")
        return int(v.value)

    def set_profiling(self, on, count_traversal: bool = False) -> None:
        level = (2 if count_traversal else 1) if on else 0
        check(self.lib.insta_set_profiling(self.h, level), "insta_set_profiling")

    def profile(self) -> dict:
        p = InstaProfile()
        check(self.lib.insta_get_profile(self.h, C.byref(p)), "insta_get_profile")
        return {name: dict(ms=p.ms[i], launches=p.launches[i], items=p.items[i])
                for i, name in enumerate(KERNEL_CLASSES)}

    def traversal_counts(self) -> dict:
        """Nearest-triangle work counted since set_profiling(True, count_traversal=True)."""
        p = InstaProfile()
        check(self.lib.insta_get_profile(self.h, C.byref(p)), "insta_get_profile")
        return dict(queries=p.bvh_queries, inner=p.bvh_inner, leaves=p.bvh_leaves, triangles=p.bvh_triangles,
                    list_queries=p.list_queries, list_entries=p.list_entries)

    # --- occupancy samples and density grid ---------------------------------
    def deform_positions(self, pos_dev, n_frames: int, seed: int, rng_offset: int = 0) -> None:
        check(self.lib.insta_deform_positions(self.h, _ptr(pos_dev), pos_dev.shape[0], n_frames, seed, rng_offset,
                                              self.stream()), "insta_deform_positions")

    def update_density_grid(self, slot: int, n_frames: int, decay: float = 0.95, n_uniform: int = 0,
                            n_nonuniform: int = 0, seed: int = 1337, want_grid: bool = False):
        """One EMA update; returns (grid [G^3] tensor or None, mean)."""
        cfg = InstaDensityGridConfig(decay, n_uniform, n_nonuniform, seed)
        G = self.cfg.grid_size
        grid = self.torch.empty(G ** 3, dtype=self.torch.float32, device="cuda") if want_grid else None
        mean = C.c_float()
        check(self.lib.insta_update_density_grid(self.h, slot, n_frames, C.byref(cfg), _ptr(grid), C.byref(mean),
                                                 self.stream()), "insta_update_density_grid")
        return grid, mean.value

    def reset_density_grid(self) -> None:
        check(self.lib.insta_reset_density_grid(self.h), "insta_reset_density_grid")

    # --- training ---------------------------------------------------------
    def set_training_set(self, images_dev, cams, frames) -> None:
        """images_dev [n, h, w, 4] float32 on the device; cams: InstaCamera per image; frames: ints."""
        n, h, w, _ = images_dev.shape
        arr = (InstaCamera * n)(*cams)
        fr = np.ascontiguousarray(frames, dtype=np.int32)
        check(self.lib.insta_set_training_set(self.h, n, w, h, _ptr(images_dev.contiguous()), C.addressof(arr),
                                              fr.ctypes.data, self.stream()), "insta_set_training_set")

    def set_training_supervision(self, seg_dev=None, depth_dev=None) -> None:
        """seg_dev [n, h, w] uint8 mask ids and depth_dev [n, h, w] float32
        camera-axis depth (0 = none) on the device, either None, for the
        training set set last."""
        check(self.lib.insta_set_training_supervision(
            self.h, None if seg_dev is None else _ptr(seg_dev.contiguous()),
            None if depth_dev is None else _ptr(depth_dev.contiguous()), self.stream()),
            "insta_set_training_supervision")

    def _reserve_training(self, tc: "InstaTrainConfig") -> None:
        # the adaptive ray count stays below 2^18 rays
        self._reserve_for(train_rays=tc.n_rays or (1 << 18), train_samples=tc.max_samples,
                          train_batch=tc.target_samples)

    def train_step(self, tc: "InstaTrainConfig") -> InstaTrainStats:
        self._reserve_training(tc)
        st = InstaTrainStats()
        check(self.lib.insta_train_step(self.h, C.byref(tc), C.byref(st), self.stream()), "insta_train_step")
        return st

    def train_loop(self, tc: "InstaTrainConfig", n_steps: int, density_slot: int, n_frames: int,
                   decay: float = 0.95, seed: int = 1337) -> InstaTrainStats:
        """n_steps of the reference's loop: density-grid EMA on its schedule into
        density_slot, which the rays march."""
        self._reserve_training(tc)
        st = InstaTrainStats()
        dg = InstaDensityGridConfig(decay, 0, 0, seed)
        check(self.lib.insta_train_loop(self.h, C.byref(tc), n_steps, C.byref(dg), density_slot, n_frames, C.byref(st),
                                        self.stream()), "insta_train_loop")
        return st

    def gradients(self):
        g = self.torch.empty(self.layout.n_params, dtype=self.torch.float32, device="cuda")
        check(self.lib.insta_get_gradients(self.h, _ptr(g), self.stream()), "insta_get_gradients")
        return g

    def params(self):
        p = self.torch.empty(self.layout.n_params, dtype=self.torch.float32, device="cuda")
        check(self.lib.insta_get_params(self.h, _ptr(p), self.stream()), "insta_get_params")
        return p

    def inference_params(self):
        p = self.torch.empty(self.layout.n_params, dtype=self.torch.float32, device="cuda")
        check(self.lib.insta_get_inference_params(self.h, _ptr(p), self.stream()), "insta_get_inference_params")
        return p

    def reset_optimizer(self) -> None:
        check(self.lib.insta_reset_optimizer(self.h, self.stream()), "insta_reset_optimizer")

    def render_stats(self) -> InstaRenderStats:
        st = InstaRenderStats()
        check(self.lib.insta_get_render_stats(self.h, C.byref(st)), "insta_get_render_stats")
        return st


def load_config_json(path) -> tuple[InstaConfig, InstaRunConfig]:
    """The reference's JSON network configuration ("parent" chain merged)."""
    cfg, run = InstaConfig(), InstaRunConfig()
    check(load_library().insta_load_config_json(str(path).encode(), C.byref(cfg), C.byref(run)),
          "insta_load_config_json")
    return cfg, run


def default_frame_config(**kw) -> InstaFrameConfig:
    fc = InstaFrameConfig()
    load_library().insta_default_frame_config(C.byref(fc))
    for k, v in kw.items():
        if k == "background":
            for i in range(4):
                fc.background[i] = float(v[i])
        else:
            setattr(fc, k, v)
    return fc


def frame_buffer(rgba, depth, spp: int, fc: InstaFrameConfig, accum, out, depth_out=None, stream=None) -> None:
    """insta_frame_buffer on device tensors (rgba [n, 4], accum/out [n, 4])."""
    import torch
    s = stream if stream is not None else C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(load_library().insta_frame_buffer(_ptr(rgba), _ptr(depth), rgba.shape[0], spp, C.byref(fc), _ptr(accum),
                                            _ptr(out), _ptr(depth_out), s), "insta_frame_buffer")


def default_train_config(**kw) -> InstaTrainConfig:
    tc = InstaTrainConfig()
    load_library().insta_default_train_config(C.byref(tc))
    for k, v in kw.items():
        setattr(tc, k, v)
    return tc


def snapshot_config(path) -> InstaConfig:
    """The configuration stored in a snapshot file (no device work)."""
    cfg = InstaConfig()
    check(load_library().insta_snapshot_config(str(path).encode(), C.byref(cfg)), "insta_snapshot_config")
    return cfg


def default_server_config(**kw) -> InstaServerConfig:
    c = InstaServerConfig()
    load_library().insta_default_server_config(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Server:
    """Request loop over resident avatars (insta_server_*)."""

    def __init__(self, cfg: InstaServerConfig | None = None):
        self.lib = load_library()
        self.cfg = cfg if cfg is not None else default_server_config()
        h = C.c_void_p()
        check(self.lib.insta_server_create(C.byref(self.cfg), C.byref(h)), "insta_server_create")
        self.h = h
        self._keep = {}     # host coefficient arrays stay alive until submit copied them

    def add_avatar(self, ctx: Context, bases_dev) -> int:
        aid = C.c_uint32()
        check(self.lib.insta_server_add_avatar(self.h, ctx.h, _ptr(bases_dev), bases_dev.shape[0], C.byref(aid)),
              "insta_server_add_avatar")
        return aid.value

    def submit(self, avatar: int, coeffs: np.ndarray, cam: InstaCamera, width: int, height: int, rgba_dev,
               tag: int = 0) -> int:
        c = np.ascontiguousarray(coeffs, dtype=np.float32)
        req = InstaRequest(avatar, c.ctypes.data, cam, width, height, _ptr(rgba_dev), tag)
        t = C.c_uint64()
        check(self.lib.insta_server_submit(self.h, C.byref(req), C.byref(t)), "insta_server_submit")
        return t.value

    def poll(self, wait: bool = False, max_results: int = 256) -> list:
        buf = (InstaResult * max_results)()
        n = C.c_uint32()
        check(self.lib.insta_server_poll(self.h, buf, max_results, C.byref(n), 1 if wait else 0), "insta_server_poll")
        return [buf[i] for i in range(n.value)]

    def clock(self) -> float:
        return self.lib.insta_server_clock(self.h)

    def graph_stats(self) -> tuple[int, int]:
        """(requests that updated a worker's graph in place, graphs instantiated)."""
        u, b = C.c_uint64(), C.c_uint64()
        check(self.lib.insta_server_graph_stats(self.h, C.byref(u), C.byref(b)), "insta_server_graph_stats")
        return u.value, b.value

    def close(self) -> None:
        if self.h:
            self.lib.insta_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
