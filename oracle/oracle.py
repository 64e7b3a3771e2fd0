"""ctypes binding of the CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, the package's smoke() check and
bench.py's cpu_baseline leg, never by the product path.  All buffers are
numpy arrays on the host.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from insta_amd.capi import InstaCamera, InstaConfig, InstaFrameConfig, InstaLayout

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_VP = C.c_void_p
_SIGS = [
    ("orc_abi_version", C.c_uint32, []),
    ("orc_create", _VP, [C.POINTER(InstaConfig)]),
    ("orc_destroy", None, [_VP]),
    ("orc_set_threads", None, [_VP, C.c_int]),
    ("orc_query_layout", C.c_int, [C.POINTER(InstaConfig), C.POINTER(InstaLayout)]),
    ("orc_set_canonical_mesh", C.c_int, [_VP, _VP, C.c_uint32, _VP, C.c_uint32, _VP]),
    ("orc_set_frame_vertices", C.c_int, [_VP, C.c_uint32, _VP]),
    ("orc_set_frame_extra_dims", C.c_int, [_VP, C.c_uint32, C.c_uint32, _VP]),
    ("orc_render_box", None, [_VP, _VP]),
    ("orc_adjacency", None, [_VP, _VP]),
    ("orc_frames", C.c_int, [_VP, C.c_int32, _VP, _VP, _VP]),
    ("orc_bvh_n_nodes", C.c_uint32, [_VP, C.c_int32]),
    ("orc_export_bvh", C.c_int, [_VP, C.c_int32, _VP, _VP]),
    ("orc_closest_triangles", C.c_int, [_VP, C.c_int32, _VP, C.c_uint32, _VP, _VP]),
    ("orc_closest_triangles_brute", C.c_int, [_VP, C.c_int32, _VP, C.c_uint32, _VP, _VP]),
    ("orc_triangle_distance_sq", C.c_float, [_VP, _VP]),
    ("orc_deform_samples", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, C.c_int32]),
    ("orc_build_occupancy", C.c_int, [_VP, C.c_int32, C.c_float, _VP]),
    ("orc_set_occupancy", C.c_int, [_VP, C.c_int32, _VP]),
    ("orc_set_params", C.c_int, [_VP, _VP, C.c_uint64]),
    ("orc_encode", C.c_int, [_VP, _VP, C.c_uint32, _VP]),
    ("orc_encode_indices", C.c_int, [_VP, _VP, C.c_uint32, _VP]),
    ("orc_encode_backward", C.c_int, [_VP, _VP, C.c_uint32, _VP, _VP]),
    ("orc_sh16", C.c_int, [_VP, C.c_uint32, _VP]),
    ("orc_network_inference", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, _VP]),
    ("orc_ld_random_val", C.c_float, [C.c_uint32, C.c_uint32, C.c_uint32]),
    ("orc_sobol", C.c_uint32, [C.c_uint32, C.c_uint32]),
    ("orc_owen_scramble", C.c_uint32, [C.c_uint32, C.c_uint32]),
    ("orc_morton3", C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
    ("orc_morton3_invert", C.c_uint32, [C.c_uint32]),
    ("orc_pixel_offset", None, [C.c_uint32, _VP]),
    ("orc_render", C.c_int, [_VP, C.POINTER(InstaCamera), C.c_uint32, C.c_uint32, C.c_int32, C.c_int32,
                             C.c_uint32, _VP, C.c_uint32, _VP, _VP, _VP, _VP]),
    ("orc_network_plugin_d", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, _VP, _VP, _VP, _VP]),
    ("orc_network_density_d", C.c_int, [_VP, _VP, C.c_uint32, _VP]),
    ("orc_network_density", C.c_int, [_VP, _VP, C.c_uint32, _VP]),
    ("orc_frame_buffer", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, C.POINTER(InstaFrameConfig), _VP, _VP, _VP]),
    ("orc_render_fp64", C.c_int, [_VP, C.POINTER(InstaCamera), C.c_uint32, C.c_uint32, C.c_int32, C.c_int32,
                                  C.c_uint32, _VP, C.c_uint32, _VP]),
    ("orc_set_training_set", C.c_int, [_VP, C.c_uint32, C.c_uint32, C.c_uint32, _VP, _VP, _VP]),
    ("orc_pcg32_next", C.c_uint32, [C.c_uint64, C.c_uint64, C.c_uint64]),
    ("orc_pcg32_draws", None, [C.c_uint64, C.c_uint64, C.c_uint32, _VP]),
    ("orc_train_grads", C.c_int, [_VP, C.c_uint32, C.c_uint32, C.c_uint64, _VP, _VP, _VP]),
    ("orc_set_training_occupancy", C.c_int, [_VP, C.c_int32]),
    ("orc_set_training_supervision", C.c_int, [_VP, _VP, _VP, C.c_double, C.c_double]),
    ("orc_deform_positions", C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]),
    ("orc_update_density_grid", C.c_int, [_VP, C.c_int32, C.c_uint32, C.c_float, C.c_uint32, C.c_uint32, C.c_uint32,
                                          _VP, _VP]),
    ("orc_reset_density_grid", None, [_VP]),
    ("orc_get_occupancy", C.c_int, [_VP, C.c_int32, _VP]),
]

_libs: dict = {}


def lib(path=None) -> C.CDLL:
    """The oracle library (default: liboracle.so next to this file; the
    pinning tests pass a build with a planted mistake)."""
    path = Path(path) if path is not None else LIB_PATH
    key = str(path)
    if key not in _libs:
        if not path.exists():
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(key)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _libs[key] = L
    return _libs[key]


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


class Oracle:
    """CPU mirror of an insta_ctx: same call sequence, host arrays."""

    def __init__(self, cfg: InstaConfig, lib_path=None):
        self.L = lib(lib_path)
        self.cfg = cfg
        self.h = self.L.orc_create(C.byref(cfg))
        if not self.h:
            raise ValueError("oracle rejected the configuration")
        self.layout = InstaLayout()
        self.L.orc_query_layout(C.byref(cfg), C.byref(self.layout))
        self.nf = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_threads(self, n: int) -> None:
        self.L.orc_set_threads(self.h, n)

    def set_canonical_mesh(self, verts, faces, masks) -> None:
        v, f, m = _c(verts, np.float32), _c(faces, np.int32), _c(masks, np.uint32)
        self.nf = len(f)
        assert self.L.orc_set_canonical_mesh(self.h, _p(v), len(v), _p(f), len(f), _p(m)) == 0

    def set_frame(self, frame: int, verts) -> None:
        v = _c(verts, np.float32)
        assert self.L.orc_set_frame_vertices(self.h, frame, _p(v)) == 0

    def set_frame_extra_dims(self, first: int, values) -> None:
        v = _c(values, np.float32)
        v = v.reshape(-1, v.shape[-1]) if v.ndim > 1 else v.reshape(1, -1)
        assert self.L.orc_set_frame_extra_dims(self.h, first, len(v), _p(v)) == 0

    def render_box(self) -> np.ndarray:
        out = np.zeros(6, np.float32)
        self.L.orc_render_box(self.h, _p(out))
        return out

    def adjacency(self) -> np.ndarray:
        out = np.zeros((self.nf, 3), np.int32)
        self.L.orc_adjacency(self.h, _p(out))
        return out

    def frames(self, frame: int):
        P = np.zeros((self.nf, 3, 4), np.float32)
        Pi = np.zeros((self.nf, 3, 4), np.float32)
        cen = np.zeros((self.nf, 3), np.float32)
        assert self.L.orc_frames(self.h, frame, _p(P), _p(Pi), _p(cen)) == 0
        return P, Pi, cen

    def export_bvh(self, frame: int):
        nn = self.L.orc_bvh_n_nodes(self.h, frame)
        nodes = np.zeros(nn * 8, np.float32)
        order = np.zeros(self.nf, np.uint32)
        assert self.L.orc_export_bvh(self.h, frame, _p(nodes), _p(order)) == 0
        return nodes, order

    def closest(self, frame: int, pts):
        p = _c(pts, np.float32)
        ids = np.zeros(len(p), np.int32)
        dist = np.zeros(len(p), np.float32)
        assert self.L.orc_closest_triangles(self.h, frame, _p(p), len(p), _p(ids), _p(dist)) == 0
        return ids, dist

    def closest_brute(self, frame: int, pts):
        p = _c(pts, np.float32)
        ids = np.zeros(len(p), np.int32)
        d2 = np.zeros(len(p), np.float32)
        assert self.L.orc_closest_triangles_brute(self.h, frame, _p(p), len(p), _p(ids), _p(d2)) == 0
        return ids, d2

    def deform(self, coords, frame: int = -1) -> np.ndarray:
        c = _c(coords, np.float32).copy()
        n, stride = c.shape
        assert self.L.orc_deform_samples(self.h, _p(c), n, stride, frame) == 0
        return c

    def build_occupancy(self, frame: int, threshold: float) -> np.ndarray:
        G = self.cfg.grid_size
        out = np.zeros(G ** 3 // 8, np.uint8)
        assert self.L.orc_build_occupancy(self.h, frame, threshold, _p(out)) == 0
        return out

    def get_occupancy(self, slot: int) -> np.ndarray:
        G = self.cfg.grid_size
        out = np.zeros(G ** 3 // 8, np.uint8)
        assert self.L.orc_get_occupancy(self.h, slot, _p(out)) == 0
        return out

    def set_occupancy(self, slot: int, bits) -> None:
        b = _c(bits, np.uint8)
        assert self.L.orc_set_occupancy(self.h, slot, _p(b)) == 0

    def set_params(self, params) -> None:
        p = _c(params, np.float32)
        assert self.L.orc_set_params(self.h, _p(p), p.size) == 0

    def encode(self, pos) -> np.ndarray:
        p = _c(pos, np.float32)
        out = np.zeros((len(p), self.cfg.n_levels * self.cfg.n_features_per_level), np.float32)
        assert self.L.orc_encode(self.h, _p(p), len(p), _p(out)) == 0
        return out

    def encode_indices(self, pos) -> np.ndarray:
        p = _c(pos, np.float32)
        out = np.zeros((len(p), self.cfg.n_levels, 8), np.uint32)
        assert self.L.orc_encode_indices(self.h, _p(p), len(p), _p(out)) == 0
        return out

    def encode_backward(self, pos, dL) -> np.ndarray:
        p = _c(pos, np.float32)
        d = _c(dL, np.float32)
        g = np.zeros(self.layout.n_grid_params, np.float64)
        assert self.L.orc_encode_backward(self.h, _p(p), len(p), _p(d), _p(g)) == 0
        return g

    def network(self, coords) -> np.ndarray:
        c = _c(coords, np.float32)
        out = np.zeros((len(c), 4), np.float32)
        assert self.L.orc_network_inference(self.h, _p(c), len(c), c.shape[1], _p(out)) == 0
        return out

    def render(self, cam: InstaCamera, width: int, height: int, frame: int, occupancy: int, spp: int = 0,
               pixels=None):
        sub = None if pixels is None else _c(pixels, np.uint32)
        n = width * height if sub is None else len(sub)
        rgba = np.zeros((n, 4), np.float32)
        depth = np.zeros(n, np.float32)
        steps = np.zeros(n, np.uint32)
        stats = np.zeros(2, np.uint64)
        rc = self.L.orc_render(self.h, C.byref(cam), width, height, frame, occupancy, spp, _p(sub),
                               0 if sub is None else len(sub), _p(rgba), _p(depth), _p(steps), _p(stats))
        assert rc == 0, rc
        return rgba, depth, steps, stats


def set_training_set(orc: Oracle, images, cams, frames) -> None:
    """images [n, h, w, 4] float32 (host), cams: list of InstaCamera, frames: ints."""
    im = _c(images, np.float32)
    n, h, w, _ = im.shape
    arr = (InstaCamera * n)(*cams)
    fr = _c(frames, np.int32)
    assert orc.L.orc_set_training_set(orc.h, n, w, h, _p(im), C.addressof(arr), _p(fr)) == 0


def set_training_supervision(orc: Oracle, seg=None, depth=None, seg_weight: float = 1.0,
                             depth_weight: float = 1.5) -> None:
    """Segmentation masks seg [n, h, w] uint8 (mask ids) and depth images
    depth [n, h, w] float32 (camera-axis depth, 0 = none) of the training set
    set last, either None; the weights of the segmentation and depth terms."""
    sg = None if seg is None else _c(seg, np.uint8)
    dp = None if depth is None else _c(depth, np.float32)
    assert orc.L.orc_set_training_supervision(orc.h, _p(sg), _p(dp), float(seg_weight), float(depth_weight)) == 0


def set_training_occupancy(orc: Oracle, slot: int) -> None:
    """Grid the training rays march: slot >= -1 for every ray, -2 each image's frame."""
    assert orc.L.orc_set_training_occupancy(orc.h, slot) == 0


def train_grads(orc: Oracle, n_rays: int, seed: int, rays_total: int = 0, want_grads: bool = True):
    """FP64 training batch: (grads [n_params] float64 or None, (reported loss, objective),
    (rays used, samples, samples composited))."""
    g = np.zeros(orc.layout.n_params, np.float64) if want_grads else None
    loss = np.zeros(2, np.float64)
    counts = np.zeros(3, np.uint32)
    rc = orc.L.orc_train_grads(orc.h, n_rays, seed, rays_total, _p(g), _p(loss), _p(counts))
    assert rc == 0, rc
    return g, (float(loss[0]), float(loss[1])), tuple(int(c) for c in counts)


def deform_positions(orc: Oracle, pos, n_frames: int, seed: int, rng_offset: int = 0) -> np.ndarray:
    """warp_random_samples on host positions [n, 3] (returns a new array)."""
    p = _c(pos, np.float32).copy()
    assert orc.L.orc_deform_positions(orc.h, _p(p), len(p), n_frames, seed, rng_offset) == 0
    return p


def update_density_grid(orc: Oracle, slot: int, n_frames: int, decay: float, n_uniform: int, n_nonuniform: int,
                        seed: int):
    """One density-grid update; returns (grid [G^3] float32, mean)."""
    G = orc.cfg.grid_size
    grid = np.zeros(G ** 3, np.float32)
    mean = np.zeros(1, np.float32)
    assert orc.L.orc_update_density_grid(orc.h, slot, n_frames, decay, n_uniform, n_nonuniform, seed, _p(grid),
                                         _p(mean)) == 0
    return grid, float(mean[0])


def pcg32_draws(initstate: int, initseq: int, n: int) -> np.ndarray:
    """First n outputs of PCG32 seeded (initstate, initseq), drawn one by one."""
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32_draws(initstate, initseq, n, _p(out))
    return out


def pcg32_next(initstate: int, initseq: int, skip: int = 0) -> int:
    """Output of PCG32 seeded (initstate, initseq) after `skip` draws."""
    return lib().orc_pcg32_next(initstate, initseq, skip)


def render_fp64(orc: Oracle, cam: InstaCamera, width: int, height: int, frame: int, occupancy: int, spp: int = 0,
                pixels=None) -> np.ndarray:
    """FP64 reference render (same samples as the FP32 march): rgba [n, 4] float64."""
    sub = None if pixels is None else _c(pixels, np.uint32)
    n = width * height if sub is None else len(sub)
    rgba = np.zeros((n, 4), np.float64)
    rc = orc.L.orc_render_fp64(orc.h, C.byref(cam), width, height, frame, occupancy, spp, _p(sub),
                               0 if sub is None else len(sub), _p(rgba))
    assert rc == 0, rc
    return rgba


def network_plugin_d(orc: "Oracle", coords, dL_dout=None):
    """NerfNetwork forward (and backward with dL_dout) in FP64 on NerfCoordinate
    samples [n, stride]: (out [n, 4], grads [n_params] or None, dinput [n, 6] or
    None) -- dinput holds d/dposition then d/d(direction input in [0,1])."""
    c = _c(coords, np.float64)
    n, stride = c.shape
    out = np.zeros((n, 4), np.float64)
    d = None if dL_dout is None else _c(dL_dout, np.float64)
    g = None if d is None else np.zeros(orc.layout.n_params, np.float64)
    di = None if d is None else np.zeros((n, 6), np.float64)
    assert orc.L.orc_network_plugin_d(orc.h, _p(c), n, stride, _p(d), _p(out), _p(g), _p(di)) == 0
    return out, g, di


def network_density(orc: "Oracle", pos) -> np.ndarray:
    """NerfNetwork::density in FP32: [n, 16] density MLP outputs."""
    p = _c(pos, np.float32)
    out = np.zeros((len(p), 16), np.float32)
    assert orc.L.orc_network_density(orc.h, _p(p), len(p), _p(out)) == 0
    return out


def network_density_d(orc: "Oracle", pos) -> np.ndarray:
    p = _c(pos, np.float64)
    out = np.zeros((len(p), 16), np.float64)
    assert orc.L.orc_network_density_d(orc.h, _p(p), len(p), _p(out)) == 0
    return out


def frame_buffer(rgba, depth, spp: int, fc: InstaFrameConfig, accum, lib_path=None):
    """Render-buffer stages (shade, accumulate, tonemap) on a composited
    image; `accum` [n, 4] float32 is updated in place.  Returns (out, depth)."""
    c = _c(rgba, np.float32).reshape(-1, 4)
    d = None if depth is None else _c(depth, np.float32)
    assert accum.dtype == np.float32 and accum.flags.c_contiguous and accum.size == c.size
    out = np.zeros_like(c)
    dout = np.zeros(len(c), np.float32)
    assert lib(lib_path).orc_frame_buffer(_p(c), _p(d), len(c), spp, C.byref(fc), _p(accum), _p(out), _p(dout)) == 0
    return out, dout


def sh16(dirs_w, lib_path=None) -> np.ndarray:
    d = _c(dirs_w, np.float32)
    out = np.zeros((len(d), 16), np.float32)
    lib(lib_path).orc_sh16(_p(d), len(d), _p(out))
    return out


def ld_random_val(index: int, seed: int, dim: int = 0, lib_path=None) -> float:
    return lib(lib_path).orc_ld_random_val(index, seed, dim)


def sobol(index: int, dim: int) -> int:
    return lib().orc_sobol(index, dim)


def owen_scramble(x: int, seed: int) -> int:
    return lib().orc_owen_scramble(x, seed)


def morton3(x: int, y: int, z: int) -> int:
    return lib().orc_morton3(x, y, z)


def morton3_invert(v: int) -> int:
    return lib().orc_morton3_invert(v)


def pixel_offset(spp: int) -> np.ndarray:
    out = np.zeros(2, np.float32)
    lib().orc_pixel_offset(spp, _p(out))
    return out


def triangle_distance_sq(tri9, p) -> float:
    t = _c(tri9, np.float32)
    q = _c(p, np.float32)
    return lib().orc_triangle_distance_sq(_p(t), _p(q))


def adam_update_d(w, g, m, v, t, lr, beta1, beta2, eps, l2_reg=0.0):
    """One Adam update in FP64 (Kingma & Ba 2015, with bias correction), the
    L2 term l2_reg * w joined to the gradient in the moment updates (tcnn's
    Adam "l2_reg", configs/nerf/main.json:21-32).  Returns (w, m, v)."""
    w, g, m, v = (np.asarray(x, np.float64) for x in (w, g, m, v))
    gg = g + l2_reg * w
    m = beta1 * m + (1.0 - beta1) * gg
    v = beta2 * v + (1.0 - beta2) * gg * gg
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def ema_update_d(a, w, decay, t):
    """The Ema optimizer's weight average in FP64 (tcnn "Ema", main.json
    decay 0.95): a_t = decay * a_{t-1} + (1 - decay) * w_t from a_0 = 0, and
    the debiased a_t / (1 - decay^t) that inference uses.  Returns (a, debiased)."""
    a = decay * np.asarray(a, np.float64) + (1.0 - decay) * np.asarray(w, np.float64)
    return a, a / (1.0 - decay ** t)
