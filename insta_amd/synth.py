"""Seeded synthetic inputs for the INSTA per-sample hot path.

Everything a workload needs is generated here, deterministically from integer
seeds, with numpy only: meshes (vertices, faces, per-triangle region mask ids),
posed frames, expression bases, cameras and network parameters.  The same arrays
feed the HIP library and the CPU oracle, so both see bit-identical inputs.

Coordinate conventions follow the reference application:
  * mesh vertices carry the +0.5 offset applied when an OBJ is loaded
    (src/tiny_mesh.cu:63-74), so a head centred at the origin sits at
    (0.5, 0.5, 0.5) inside the unit cube that the NeRF occupies;
  * y is up, z points out of the face ("front");
  * a camera is a 3x4 matrix whose columns are right, down, forward and the
    camera position (common_device.cuh:260-315 multiplies (u, v, 1) by the
    3x3 block and takes column 3 as the origin).

Region mask ids use the reference's FlameMasks enum values
(include/rta/masking.h:32-50): Boundary=0, EyeRegion=1, Face=2, Lips=7,
Torso=15.  Synthetic meshes carry them directly instead of parsing mask OBJs.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK_BOUNDARY = 0
MASK_EYE_REGION = 1
MASK_FACE = 2
MASK_LIPS = 7
MASK_TORSO = 15

CENTER = np.array([0.5, 0.5, 0.5], dtype=np.float64)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


# ----------------------------------------------------------------------------
# meshes
# ----------------------------------------------------------------------------

def octasphere(level: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit sphere from an octahedron subdivided `level` times (8 * 4**level faces).

    Level 4 gives the 2,048-triangle, 1,026-vertex sphere of the small workload.
    Faces are wound counter-clockwise seen from outside.
    """
    verts = [np.array(v, dtype=np.float64) for v in
             [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]]
    faces = [(4, 0, 2), (4, 2, 1), (4, 1, 3), (4, 3, 0),
             (5, 2, 0), (5, 1, 2), (5, 3, 1), (5, 0, 3)]
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def mid(i: int, j: int) -> int:
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
        faces = nf
    v = np.stack(verts)
    f = np.asarray(faces, dtype=np.int32)
    # orient every face outward
    cen = v[f].mean(axis=1)
    nrm = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    flip = (nrm * cen).sum(axis=1) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def head_ellipsoid(rings: int = 81, per_ring: int = 62,
                   theta_max: float = 0.85 * math.pi) -> tuple[np.ndarray, np.ndarray]:
    """Unit-scale ellipsoid topology with an open neck and a small mouth slit.

    A top pole plus `rings` latitude rings of `per_ring` vertices, open below
    the last ring (the neck).  With the defaults this is 1 + 81*62 = 5,023
    vertices and 62 + 2*62*80 = 9,982 triangles; removing three quads (six
    triangles) in the lower front (a mouth slit) gives 9,976 triangles, the
    vertex/triangle counts of the FLAME head mesh, without leaving any vertex
    unreferenced.  Returns directions on the unit sphere (scaled by the
    caller) and faces.
    """
    verts = [np.array([0.0, 1.0, 0.0])]
    for r in range(1, rings + 1):
        th = theta_max * r / rings
        off = (r % 2) * math.pi / per_ring
        for j in range(per_ring):
            ph = 2 * math.pi * j / per_ring + off
            verts.append(np.array([math.sin(th) * math.cos(ph), math.cos(th),
                                   math.sin(th) * math.sin(ph)]))
    v = np.stack(verts)

    def vid(r: int, j: int) -> int:
        return 1 + (r - 1) * per_ring + (j % per_ring)

    faces = []
    for j in range(per_ring):
        faces.append((0, vid(1, j + 1), vid(1, j)))
    mouth_ring = int(round(0.55 * rings))
    # ring vertex j sits at angle 2*pi*j/M (+ half-step offset on odd rings); the
    # front (+z) is at angle pi/2
    j_front = int(round(per_ring / 4))
    mouth = {(mouth_ring, j_front - 1), (mouth_ring, j_front), (mouth_ring, j_front + 1)}
    for r in range(1, rings):
        for j in range(per_ring):
            if (r, j) in mouth:
                continue
            a, b = vid(r, j), vid(r, j + 1)
            c, d = vid(r + 1, j), vid(r + 1, j + 1)
            faces.append((a, b, d))
            faces.append((a, d, c))
    f = np.asarray(faces, dtype=np.int32)
    cen = v[f].mean(axis=1)
    nrm = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    flip = (nrm * cen).sum(axis=1) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def region_masks(verts: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """Per-triangle FlameMasks ids from the triangle centroid direction.

    Disjoint patches: lips (7) in the lower front, two eye regions (1), the rest
    of the front is face (2), a bottom cap is boundary (0), everything else
    torso (15).  Face triangles take the interpolated-frame deformation path,
    eye-region triangles the area-scaled frame, lips keep expression dims.
    """
    cen = verts[faces].mean(axis=1) - CENTER
    u = cen / np.linalg.norm(cen, axis=1, keepdims=True)
    x, y, z = u[:, 0], u[:, 1], u[:, 2]
    m = np.full(len(faces), MASK_TORSO, dtype=np.uint32)
    eye = (z > 0.45) & (y > 0.1) & (y < 0.45) & (np.abs(x) > 0.12) & (np.abs(x) < 0.55)
    lips = (z > 0.55) & (y > -0.5) & (y < -0.2) & (np.abs(x) < 0.35)
    face = (z > 0.2) & ~eye & ~lips
    m[face] = MASK_FACE
    m[eye] = MASK_EYE_REGION
    m[lips] = MASK_LIPS
    m[y < -0.9] = MASK_BOUNDARY
    return m


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform rotation on SO(3) from a normalised Gaussian quaternion."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


@dataclass
class Avatar:
    """A canonical mesh, its posed frames and the optional expression bases."""
    canonical: np.ndarray            # [V, 3] float32
    faces: np.ndarray                # [F, 3] int32
    masks: np.ndarray                # [F] uint32
    frames: list = field(default_factory=list)   # list of [V, 3] float32
    bases: np.ndarray | None = None  # [K, V, 3] float32 displacement bases


def small_sphere_avatar(seed: int = 0, n_frames: int = 1, n_bases: int = 0) -> Avatar:
    """Small workload mesh: 2,048-triangle sphere of radius 0.3, N(0, 0.005)
    vertex noise; each posed frame is a uniform random rotation about the
    centre plus per-vertex N(0, 0.01) displacement.  With n_bases > 0 the
    avatar also gets that many Gaussian expression bases (sigma 3 mm, width
    5 cm) for per-request expressions."""
    rng = _rng(seed)
    v, f = octasphere(4)
    v = v * 0.3 + rng.normal(0.0, 0.005, size=v.shape)
    canon = (v + CENTER).astype(np.float32)
    masks = region_masks(canon.astype(np.float64), f)
    frames = []
    for _ in range(n_frames):
        rot = random_rotation(rng)
        p = (canon.astype(np.float64) - CENTER) @ rot.T + CENTER
        p = p + rng.normal(0.0, 0.01, size=p.shape)
        frames.append(p.astype(np.float32))
    bases = gaussian_bases(canon, rng, n_bases=n_bases, width=0.05) if n_bases else None
    return Avatar(canon, f, masks, frames, bases)


def blend_expression_f32(canon: np.ndarray, bases: np.ndarray, coeffs: np.ndarray) -> np.ndarray:
    """canonical + sum_k coeffs[k] * bases[k] accumulated in float32 in k
    order, one rounding per multiply and per add (the device blend's order)."""
    acc = canon.astype(np.float32).copy()
    c = np.asarray(coeffs, np.float32)
    for k in range(len(c)):
        acc = (acc + c[k] * bases[k]).astype(np.float32)
    return acc


def _low_frequency_field(u: np.ndarray, rng: np.random.Generator, amp: float) -> np.ndarray:
    """Smooth scalar field on the unit sphere: random combination of degree<=3
    monomials of the direction, scaled to a peak of about `amp`."""
    x, y, z = u[:, 0], u[:, 1], u[:, 2]
    terms = [x, y, z, x * y, y * z, x * z, x * x - y * y, 3 * z * z - 1,
             x * x * x, y * y * y, z * z * z, x * y * z]
    c = rng.normal(0.0, 1.0, size=len(terms))
    s = sum(ci * t for ci, t in zip(c, terms))
    return amp * s / max(np.abs(s).max(), 1e-12)


def gaussian_bases(canon: np.ndarray, rng: np.random.Generator, n_bases: int = 50,
                   sigma: float = 0.003, width: float = 0.025) -> np.ndarray:
    """`n_bases` smooth per-vertex displacement fields: a Gaussian bump of
    spatial width `width` around a random vertex, displacing along a random
    N(0, sigma^2 I) vector."""
    v = canon.astype(np.float64)
    centres = rng.integers(0, len(v), size=n_bases)
    dirs = rng.normal(0.0, sigma, size=(n_bases, 3))
    out = np.empty((n_bases, len(v), 3), dtype=np.float32)
    for k in range(n_bases):
        d2 = ((v - v[centres[k]]) ** 2).sum(axis=1)
        g = np.exp(-d2 / (2 * width * width))
        out[k] = (g[:, None] * dirs[k][None, :]).astype(np.float32)
    return out


def blend_expression(canon: np.ndarray, bases: np.ndarray, coeffs: np.ndarray) -> np.ndarray:
    """Vertices of one expression: canonical + sum_k coeffs[k] * bases[k]."""
    d = np.tensordot(coeffs.astype(np.float64), bases.astype(np.float64), axes=(0, 0))
    return (canon.astype(np.float64) + d).astype(np.float32)


def head_avatar(seed: int = 1, n_frames: int = 16, n_bases: int = 50) -> Avatar:
    """Full-size workload mesh: 5,023-vertex / 9,976-triangle ellipsoid with
    semi-axes 0.09 / 0.12 / 0.11 m (x / y / z) and a seeded low-frequency
    displacement of up to 5 mm along the radial direction; `n_frames` posed
    frames, each a N(0, 1) blend of `n_bases` Gaussian displacement bases
    (sigma 3 mm)."""
    rng = _rng(seed)
    u, f = head_ellipsoid()
    axes = np.array([0.09, 0.12, 0.11])
    disp = _low_frequency_field(u, rng, 0.005)
    v = u * axes * (1.0 + disp[:, None] / 0.1)
    canon = (v + CENTER).astype(np.float32)
    masks = region_masks(canon.astype(np.float64), f)
    bases = gaussian_bases(canon, rng, n_bases=n_bases)
    frames = [blend_expression(canon, bases, rng.standard_normal(n_bases)) for _ in range(n_frames)]
    return Avatar(canon, f, masks, frames, bases)


# ----------------------------------------------------------------------------
# cameras
# ----------------------------------------------------------------------------

def look_at_camera(direction: np.ndarray, radius: float, target=CENTER) -> np.ndarray:
    """3x4 float32 camera matrix [right | down | forward | position] looking at
    `target` from `target + radius * direction` (y is world up)."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    pos = np.asarray(target, dtype=np.float64) + radius * d
    fwd = -d
    up = np.array([0.0, 1.0, 0.0])
    if abs(np.dot(fwd, up)) > 0.999:
        up = np.array([0.0, 0.0, 1.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    return np.stack([right, down, fwd, pos], axis=1).astype(np.float32)


def random_camera(rng: np.random.Generator, radius: float, front_bias: bool = True) -> np.ndarray:
    """Camera on a sphere of `radius` around the centre.  With `front_bias`
    the direction is drawn from the +z hemisphere (the face side)."""
    d = rng.standard_normal(3)
    if front_bias:
        d[2] = abs(d[2]) + 0.5
    return look_at_camera(d, radius)


def focal_from_fov(height: int, fov_deg: float) -> float:
    """Focal length in pixels for a vertical field of view."""
    return 0.5 * height / math.tan(math.radians(fov_deg) * 0.5)


# ----------------------------------------------------------------------------
# training targets
# ----------------------------------------------------------------------------

def smooth_noise_images(seed: int, n: int, height: int, width: int, n_waves: int = 6) -> np.ndarray:
    """Training targets: RGBA float32 [n, height, width, 4], alpha 1, each
    colour channel a seeded smooth field in [0, 1]:
    0.5 + 0.5 * sum_k a_k sin(2 pi f_k u + p_k) cos(2 pi g_k v + q_k) / sum_k |a_k|
    over pixel-centre coordinates u, v in (0, 1), with 1-4 cycles per image."""
    rng = _rng(seed)
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    out = np.ones((n, height, width, 4), dtype=np.float32)
    for i in range(n):
        for c in range(3):
            a = rng.uniform(0.2, 1.0, n_waves)
            f = rng.uniform(1.0, 4.0, n_waves)
            g = rng.uniform(1.0, 4.0, n_waves)
            ph = rng.uniform(0.0, 2 * math.pi, (2, n_waves))
            sx = np.sin(2 * math.pi * f[:, None] * u[None, :] + ph[0][:, None])   # [K, W]
            cy = np.cos(2 * math.pi * g[:, None] * v[None, :] + ph[1][:, None])   # [K, H]
            field = np.einsum("k,kh,kw->hw", a, cy, sx) / a.sum()
            out[i, :, :, c] = (0.5 + 0.5 * field).astype(np.float32)
    return out


# ----------------------------------------------------------------------------
# network parameters (declared order: density MLP, colour MLP, hash tables)
# ----------------------------------------------------------------------------

def xavier_uniform(rng: np.random.Generator, fan_out: int, fan_in: int) -> np.ndarray:
    a = math.sqrt(6.0 / (fan_in + fan_out))
    return rng.uniform(-a, a, size=(fan_out, fan_in)).astype(np.float32)


def network_params(rng_seed: int, layout, cfg) -> np.ndarray:
    """Flat FP32 parameter vector for `layout` (an insta_amd.capi.InstaLayout
    of `cfg`): Xavier-uniform bias-free MLP weights, hash-grid features
    U(-1e-4, 1e-4)."""
    rng = _rng(rng_seed)
    parts = []
    for (fan_out, fan_in) in layout.mlp_shapes(cfg):
        parts.append(xavier_uniform(rng, fan_out, fan_in).ravel())
    parts.append(rng.uniform(-1e-4, 1e-4, size=layout.n_grid_params).astype(np.float32))
    out = np.concatenate(parts)
    assert out.size == layout.n_params
    return out
