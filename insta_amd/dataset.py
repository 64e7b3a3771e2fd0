"""Dataset ingestion: an INSTA capture (transforms JSON, per-frame tracked
meshes as OBJ, per-frame expression files, images) turned into the inputs of
insta_set_canonical_mesh, insta_set_frames and insta_set_training_set.

Follows the reference's loaders for the parts this path uses:
  * load_nerf (instant-ngp/src/nerf_loader.cu:405-900): every frame list is
    sorted by "mesh_path" (433-435), "n_frames" keeps the first n (437-440),
    "scale" / "offset" / "aabb" set the scene transform (612-636), the focal
    length comes from "x_fov" (degrees), "fl_x" or "camera_angle_x" (radians)
    at the file level, then the frame level (339-366, 863-864), the principal
    point from cx / w and cy / h (290-296, 324-337), and a camera matrix is
    the frame's "transform_matrix" with its translation moved by the offset
    (mica_matrix_to_ngp, nerf_loader.h:138-143: MICA cameras need no axis
    change; their columns are right, down, forward, position);
  * TinyMesh::load (src/tiny_mesh.cu:53-87): OBJ vertices + 0.5 per axis
    (MESH_SCALE = 1), faces as vertex triples (polygons fan-triangulated, as
    tinyobjloader does by default);
  * read_flame_params (include/rta/helpers.h:30-46): one float per line, at
    most 75;
  * Core::load_meshes (src/core.cu:234-337): the canonical mesh is
    "canonical.obj" next to the transforms file;
  * rta::Masking (src/masking.cpp:120-178): FLAME region masks from the
    vertex colours of mesh_masks/{boundary, eye_region, face, left_ear, lips,
    right_ear}.obj (a vertex is in a region when its colour is (>= 1, 0, 0));
    a face takes the region of the highest FlameMasks id among the regions
    that hold one of its vertices, else TORSO.

  * segmentation masks and depth images (nerf_loader.cu:773-823, 992-1012):
    a frame's "seg_mask_path" image (its first channel is the mask id); with
    "integer_depth_scale" > 0 at the file level, a frame's "depth_path" image
    (16-bit PNG, or .bin: one uint32 per pixel kept as uint16) times that
    scale times the scene scale; 0 where a frame has none.

Not covered: camera distortion and rolling shutter, EXR images, alpha /
dynamic-mask side images, ray files and dataset caching.  The shuffled
subset of training images and its periodic reload are insta_amd.training's.  Without a
mesh_masks folder every face is TORSO (the plain per-triangle frame) unless
per-face ids are passed.  Images are returned as stored (value / 255), with
no sRGB-to-linear conversion, as the training step takes its targets.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .capi import InstaCamera
from .synth import MASK_TORSO

# mesh_masks files and their FlameMasks ids (include/rta/masking.h:32-50)
MESH_MASK_REGIONS = (("boundary", 0), ("eye_region", 1), ("face", 2), ("left_ear", 4), ("lips", 7),
                     ("right_ear", 10))

MESH_SCALE = 1.0
MESH_OFFSET = 0.5
N_FLAME_PARAMS = 75


def read_obj(path) -> tuple[np.ndarray, np.ndarray]:
    """(vertices [V, 3] float32 with the +0.5 offset, faces [F, 3] int32)."""
    verts, faces = [], []
    for line in Path(path).read_text().splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            verts.append([float(parts[1]), float(parts[2]), float(parts[3])])
        elif parts[0] == "f":
            idx = []
            for tok in parts[1:]:
                k = int(tok.split("/")[0])
                idx.append(k - 1 if k > 0 else len(verts) + k)  # 1-based, or relative when negative
            for j in range(1, len(idx) - 1):
                faces.append([idx[0], idx[j], idx[j + 1]])
    v = np.asarray(verts, dtype=np.float32).reshape(-1, 3)
    f = np.asarray(faces, dtype=np.int32).reshape(-1, 3)
    if len(f) and (f.min() < 0 or f.max() >= len(v)):
        raise ValueError(f"{path}: face index out of range")
    return (v * np.float32(MESH_SCALE) + np.float32(MESH_OFFSET)).astype(np.float32), f


def read_mask_vertices(path) -> tuple[np.ndarray, int]:
    """(ids of the vertices coloured (>= 1, 0, 0), vertex count) of an OBJ
    whose "v" lines carry x y z r g b (Masking::load_mesh_color_mask); a
    vertex without a colour is not in the region."""
    ids, n = [], 0
    for line in Path(path).read_text().splitlines():
        parts = line.split()
        if not parts or parts[0] != "v":
            continue
        if len(parts) >= 7:
            r, g, b = float(parts[4]), float(parts[5]), float(parts[6])
            if r >= 1.0 and g == 0.0 and b == 0.0:
                ids.append(n)
        n += 1
    return np.asarray(ids, dtype=np.int64), n


def region_masks(faces: np.ndarray, n_vertices: int, mesh_masks_dir) -> np.ndarray:
    """Per-face FlameMasks ids from a mesh_masks folder (Masking::mesh_to_mask
    and triangle_to_mask): the regions are visited in id order and each one
    that holds a vertex of the face overwrites its id, so the highest wins."""
    faces = np.asarray(faces, dtype=np.int64)
    out = np.full(len(faces), MASK_TORSO, np.uint32)
    for name, rid in MESH_MASK_REGIONS:
        ids, n = read_mask_vertices(Path(mesh_masks_dir) / f"{name}.obj")
        if n != n_vertices:
            raise ValueError(f"{name}.obj: {n} vertices, the canonical mesh has {n_vertices}")
        member = np.zeros(n_vertices, bool)
        member[ids] = True
        out[member[faces].any(axis=1)] = rid
    return out


def read_flame_params(path, max_params: int = N_FLAME_PARAMS) -> np.ndarray:
    out = []
    for line in Path(path).read_text().splitlines():
        if line and len(out) < max_params:
            out.append(float(line))
    return np.asarray(out, dtype=np.float32)


def _focal(doc: dict, res: tuple[int, int]) -> tuple[float, float] | None:
    def axis(resolution: int, a: str) -> float:
        if a + "_fov" in doc:
            return 0.5 * resolution / math.tan(0.5 * float(doc[a + "_fov"]) * math.pi / 180.0)
        if "fl_" + a in doc:
            return float(doc["fl_" + a])
        if "camera_angle_" + a in doc:
            deg = float(doc["camera_angle_" + a]) * 180.0 / math.pi
            return 0.5 * resolution / math.tan(0.5 * deg * math.pi / 180.0)
        return 0.0
    fx, fy = axis(res[0], "x"), axis(res[1], "y")
    if fx != 0.0:
        return fx, (fy if fy != 0.0 else fx)
    if fy != 0.0:
        return fy, fy
    return None


def _principal_point(doc: dict, default: tuple[float, float]) -> tuple[float, float]:
    px, py = default
    if "cx" in doc and "w" in doc:
        px = float(doc["cx"]) / float(doc["w"])
    if "cy" in doc and "h" in doc:
        py = float(doc["cy"]) / float(doc["h"])
    return px, py


@dataclass
class Dataset:
    canonical: np.ndarray              # [V, 3] float32
    faces: np.ndarray                  # [F, 3] int32
    masks: np.ndarray                  # [F] uint32
    frames: list = field(default_factory=list)        # per image: posed mesh [V, 3] float32
    expressions: list = field(default_factory=list)   # per image: FLAME parameters (float32)
    cameras: list = field(default_factory=list)       # per image: InstaCamera
    image_paths: list = field(default_factory=list)
    seg_paths: list = field(default_factory=list)     # per image: mask image path or None
    depth_paths: list = field(default_factory=list)   # per image: depth image path or None
    depth_scales: list = field(default_factory=list)  # per image: metres-to-scene factor, < 0: no depth
    file_counts: list = field(default_factory=list)   # images per transforms file, in order
    resolution: tuple = (0, 0)         # (width, height)
    offset: tuple = (0.5, 0.5, 0.5)
    scale: float = 1.0

    def _ids(self, ids):
        return range(len(self.image_paths)) if ids is None else list(ids)

    def images(self, ids=None) -> np.ndarray:
        """[n, h, w, 4] float32 RGBA in [0, 1] of images `ids` (default all;
        PIL decodes the files)."""
        from PIL import Image
        w, h = self.resolution
        ids = self._ids(ids)
        out = np.empty((len(ids), h, w, 4), np.float32)
        for k, i in enumerate(ids):
            p = self.image_paths[i]
            with Image.open(p) as im:
                a = np.asarray(im.convert("RGBA"), dtype=np.float32)
            if a.shape[:2] != (h, w):
                raise ValueError(f"{p}: {a.shape[1]}x{a.shape[0]}, expected {w}x{h}")
            out[k] = a / 255.0
        return out

    def seg_masks(self, ids=None) -> np.ndarray | None:
        """[n, h, w] uint8 mask ids of images `ids` (the first channel of each
        mask image), None when none of them has a mask.  The device takes
        masks for every image of a training set or for none."""
        from PIL import Image
        ids = self._ids(ids)
        have = [self.seg_paths[i] is not None for i in ids]
        if not any(have):
            return None
        if not all(have):
            raise ValueError("segmentation masks for some images of the set only")
        w, h = self.resolution
        out = np.empty((len(ids), h, w), np.uint8)
        for k, i in enumerate(ids):
            with Image.open(self.seg_paths[i]) as im:
                a = np.asarray(im.convert("RGBA"), dtype=np.uint8)
            if a.shape[:2] != (h, w):
                raise ValueError(f"{self.seg_paths[i]}: wrong resolution")
            out[k] = a[..., 0]
        return out

    def depth_images(self, ids=None) -> np.ndarray | None:
        """[n, h, w] float32 camera-axis depth in scene units of images `ids`
        (raw value x integer_depth_scale x scene scale), 0 where an image has
        no depth file; None when no image's file sets integer_depth_scale."""
        ids = self._ids(ids)
        if not any(self.depth_scales[i] >= 0.0 for i in ids):
            return None
        w, h = self.resolution
        out = np.zeros((len(ids), h, w), np.float32)
        for k, i in enumerate(ids):
            if self.depth_paths[i] is not None and self.depth_scales[i] > 0.0:
                out[k] = read_depth(self.depth_paths[i], w, h).astype(np.float32) * np.float32(self.depth_scales[i])
        return out


def read_depth(path, width: int, height: int) -> np.ndarray:
    """[h, w] uint16 raw depth: a .bin file of one little-endian uint32 per
    pixel kept modulo 2^16 (nerf_loader.cu:794-806), else an image read as
    one 16-bit channel (stbi_load_16: 8-bit values v become 257 v)."""
    path = Path(path)
    if ".bin" in path.name:
        raw = np.fromfile(path, dtype="<u4", count=width * height)
        if raw.size != width * height:
            raise ValueError(f"{path}: expected {width * height} pixels")
        return (raw & 0xFFFF).astype(np.uint16).reshape(height, width)
    from PIL import Image
    with Image.open(path) as im:
        if im.mode in ("I;16", "I;16L", "I;16B", "I"):
            a = np.asarray(im).astype(np.int64)
        elif im.mode == "L":
            a = np.asarray(im).astype(np.int64) * 257
        else:
            raise ValueError(f"{path}: depth images are one 8- or 16-bit channel, not {im.mode}")
    if a.shape != (height, width):
        raise ValueError(f"{path}: depth image has the wrong resolution")
    return np.clip(a, 0, 65535).astype(np.uint16)


def _image_size(path: Path) -> tuple[int, int]:
    from PIL import Image
    with Image.open(path) as im:
        return im.size


def load_transforms(paths, masks: np.ndarray | None = None, mesh_masks_dir=None) -> Dataset:
    """Read one or more transforms JSON files of one capture (see the module
    docstring for the keys honoured).  Frame i of the result is image i,
    posed mesh i and expression i.  Region masks: `masks` (one id per face),
    else those of `mesh_masks_dir`, else TORSO everywhere."""
    paths = [Path(p) for p in ([paths] if isinstance(paths, (str, Path)) else paths)]
    base0 = paths[0].parent
    canon, faces = read_obj(base0 / "canonical.obj")
    if masks is None and mesh_masks_dir is not None:
        masks = region_masks(faces, len(canon), mesh_masks_dir)
    if masks is None:
        masks = np.full(len(faces), MASK_TORSO, np.uint32)
    masks = np.asarray(masks, dtype=np.uint32)
    if masks.shape != (len(faces),):
        raise ValueError("masks: one id per face of canonical.obj")
    ds = Dataset(canon, faces, masks)
    for path in paths:
        doc = json.loads(path.read_text())
        base = path.parent
        frames = doc.get("frames")
        if not isinstance(frames, list):
            continue
        frames = sorted(frames, key=lambda fr: fr["mesh_path"])
        if "n_frames" in doc:
            frames = frames[:min(len(frames), int(doc["n_frames"]))]
        scale, offset = 1.0, np.array([0.5, 0.5, 0.5])
        if "scale" in doc:
            scale = float(doc["scale"])
        if "offset" in doc:
            o = doc["offset"]
            offset = np.array([float(x) for x in o]) if isinstance(o, list) else np.full(3, float(o))
        if "aabb" in doc:
            lo, hi = np.array(doc["aabb"][0], float), np.array(doc["aabb"][1], float)
            length = max(1e-6, float(np.abs(hi - lo).max()))
            scale = 1.0 / length
            offset = (hi + lo) * 0.5 * -scale + 0.5
        ds.scale, ds.offset = scale, tuple(float(x) for x in offset)
        pp_file = _principal_point(doc, (0.5, 0.5))
        # the reference multiplies by the scene scale (m.depth_scale * result.scale)
        depth_scale = float(doc["integer_depth_scale"]) * scale if "integer_depth_scale" in doc else -1.0
        depth_on = bool(doc.get("enable_depth_loading", True))
        n_before = len(ds.cameras)
        for fr in frames:
            img = base / fr["file_path"]
            if img.suffix == "":
                img = img.with_suffix(".png")
            res = _image_size(img)
            if ds.resolution == (0, 0):
                ds.resolution = res
            elif res != ds.resolution:
                raise ValueError(f"{img}: every image of a training set has one resolution")
            focal = _focal(fr, res) or _focal(doc, res)
            if focal is None:
                raise ValueError(f"{img}: no focal length (x_fov, fl_x or camera_angle_x)")
            m = np.asarray(fr["transform_matrix"], dtype=np.float64)[:3, :4].copy()
            m[:, 3] = m[:, 3] * MESH_SCALE + offset
            cam = InstaCamera()
            for i, v in enumerate(m.astype(np.float32).ravel()):
                cam.matrix[i] = float(v)
            cam.focal_px[0], cam.focal_px[1] = focal
            cam.screen_center[0], cam.screen_center[1] = _principal_point(fr, pp_file)
            cam.snap_to_pixel_centers = 0
            verts, f = read_obj(base / fr["mesh_path"])
            if verts.shape != canon.shape or not np.array_equal(f, faces):
                raise ValueError(f"{fr['mesh_path']}: topology differs from canonical.obj")
            ds.frames.append(verts)
            ds.expressions.append(read_flame_params(base / fr["exp_path"]) if "exp_path" in fr
                                  else np.zeros(0, np.float32))
            ds.cameras.append(cam)
            ds.image_paths.append(img)
            seg = base / fr["seg_mask_path"] if "seg_mask_path" in fr else None
            ds.seg_paths.append(seg if seg is not None and seg.exists() else None)
            dp = base / fr["depth_path"] if "depth_path" in fr and depth_on and depth_scale > 0.0 else None
            ds.depth_paths.append(dp if dp is not None and dp.exists() else None)
            ds.depth_scales.append(depth_scale)
        ds.file_counts.append(len(ds.cameras) - n_before)
    if not ds.cameras:
        raise ValueError("no frames found in the transforms files")
    return ds
