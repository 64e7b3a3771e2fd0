"""Writes tests/golden/small_scalar.npz: the oracle's outputs for the small
scalar workload (a 2,048-triangle sphere, one posed frame, L=8 F=2 T=2^14
hash grid, 64^3 occupancy at distance 0.05, 64x64 render from a camera at
radius 1.0 with a 30 degree field of view, at most 256 steps of 1/256).

Inputs are regenerated from seeds by insta_amd.synth; the file keeps a
SHA-256 digest of them so that a change of the generators shows as such.
Outputs: nearest-triangle ids and distances of 512 points, 256 deformed
samples, the occupancy bitfield, hash-grid features and table indices of
128 points, network outputs of 256 samples, and the render (RGBA, depth,
step counts).

Run from the repository root with the oracle built (make -C oracle):
    python tests/golden/make_small_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

OUT = Path(__file__).resolve().parent / "small_scalar.npz"
AVATAR_SEED, PARAMS_SEED, POINTS_SEED, CAMERA_SEED = 0, 7, 123, 5
THRESHOLD, RES = 0.05, 64


def config():
    from insta_amd.capi import small_config
    return small_config(max_frames=1)


def inputs(cfg, layout):
    """The workload's inputs and probe points, from their seeds (layout: the
    parameter layout of cfg)."""
    from insta_amd import synth
    from insta_amd.capi import make_camera
    av = synth.small_sphere_avatar(seed=AVATAR_SEED, n_frames=1)
    params = synth.network_params(PARAMS_SEED, layout, cfg)
    rng = np.random.default_rng(POINTS_SEED)
    frame = av.frames[0].astype(np.float32)
    tri = frame[av.faces].astype(np.float64)
    k = rng.integers(0, len(av.faces), 384)
    bary = rng.dirichlet([1, 1, 1], 384)
    near = (np.einsum("nk,nkd->nd", bary, tri[k]) + 0.02 * rng.standard_normal((384, 3))).astype(np.float32)
    pts = np.concatenate([near, rng.uniform(0.0, 1.0, (128, 3)).astype(np.float32)])
    d = rng.standard_normal((256, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    coords = np.zeros((256, 11), np.float32)
    coords[:, 0:3] = pts[:256]
    coords[:, 5:8] = pts[:256]
    coords[:, 8:11] = (d + 1.0) * 0.5
    enc_pts = rng.uniform(0.02, 0.98, (128, 3)).astype(np.float32)
    cam = make_camera(synth.random_camera(np.random.default_rng(CAMERA_SEED), 1.0), synth.focal_from_fov(RES, 30.0))
    return av, params, pts, coords, enc_pts, cam


def digest(av, params, pts, coords, enc_pts, cam) -> str:
    h = hashlib.sha256()
    for a in (av.canonical, av.faces, av.masks, av.frames[0], params, pts, coords, enc_pts):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(bytes(cam))
    return h.hexdigest()


def main():
    from oracle.oracle import Oracle
    cfg = config()
    orc = Oracle(cfg)
    av, params, pts, coords, enc_pts, cam = inputs(cfg, orc.layout)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    orc.set_frame(0, av.frames[0])
    orc.set_params(params)
    ids, d2 = orc.closest_brute(0, pts)
    deformed = orc.deform(coords, 0)
    bits = orc.build_occupancy(0, THRESHOLD)
    rgba, depth, steps, _ = orc.render(cam, RES, RES, 0, 0)
    np.savez_compressed(
        OUT, digest=np.array(digest(av, params, pts, coords, enc_pts, cam)),
        closest_ids=ids.astype(np.int32), closest_dist=np.sqrt(d2).astype(np.float32),
        deformed=deformed.astype(np.float32), occupancy=bits.astype(np.uint8),
        encode=orc.encode(enc_pts), encode_indices=orc.encode_indices(enc_pts),
        network=orc.network(coords), rgba=rgba.astype(np.float32), depth=depth.astype(np.float32),
        steps=steps.astype(np.uint32))
    print("wrote", OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
