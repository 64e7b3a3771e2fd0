"""Pins the CPU oracle against closed forms, invariants and brute force.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a
swapped index, a wrong sign or order) fails it; none re-types the oracle's
formula.  CPU only.
"""
import math

import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import make_camera, small_config
from oracle import oracle as O
from tests.conftest import assert_rel_close, ensure_oracle

ensure_oracle()


# ---------------------------------------------------------------- geometry
TRI = np.array([[0.2, 0.3, 0.4], [0.6, 0.35, 0.42], [0.3, 0.7, 0.45]], np.float64)


def _tri_normal(t):
    n = np.cross(t[1] - t[0], t[2] - t[0])
    return n / np.linalg.norm(n)


def test_distance_closed_forms():
    tri9 = TRI.astype(np.float32).ravel()
    n = _tri_normal(TRI)
    c = TRI.mean(axis=0)
    # on a vertex: exactly zero
    assert O.triangle_distance_sq(tri9, TRI[1]) == 0.0
    # above the centroid at height h: h^2 (face region)
    for h in (1e-3, 0.05, 0.3):
        p = c + h * n
        assert math.isclose(O.triangle_distance_sq(tri9, p), h * h, rel_tol=2e-5)
    # beyond vertex a, away from the triangle: |p - a|^2 (vertex region)
    out = TRI[0] + 0.1 * (TRI[0] - c) / np.linalg.norm(TRI[0] - c)
    assert math.isclose(O.triangle_distance_sq(tri9, out), 0.01, rel_tol=2e-5)
    # beyond the midpoint of edge ab, in the triangle plane: squared distance to the line ab
    m = 0.5 * (TRI[0] + TRI[1])
    e = TRI[1] - TRI[0]
    perp = np.cross(n, e)
    perp /= np.linalg.norm(perp)
    if np.dot(perp, c - m) > 0:
        perp = -perp
    p = m + 0.07 * perp + 0.02 * n
    assert math.isclose(O.triangle_distance_sq(tri9, p), 0.07 ** 2 + 0.02 ** 2, rel_tol=2e-5)


@pytest.mark.parametrize("frame", [-1, 0, 1])
def test_bvh_closest_matches_brute_force_small(small_oracle, frame):
    rng = np.random.default_rng(3 + frame)
    pts = np.concatenate([rng.uniform(0.0, 1.0, (3000, 3)),
                          0.5 + 0.31 * rng.standard_normal((3000, 3)) / 1.7]).astype(np.float32)
    ids, dist = small_oracle.closest(frame, pts)
    bid, bd2 = small_oracle.closest_brute(frame, pts)
    assert np.array_equal(ids, bid)
    assert np.array_equal(dist, np.sqrt(bd2).astype(np.float32))


def test_bvh_closest_matches_brute_force_head(head_avatar):
    from oracle.oracle import Oracle
    from insta_amd.capi import full_config
    o = Oracle(full_config(max_frames=4))
    o.set_canonical_mesh(head_avatar.canonical, head_avatar.faces, head_avatar.masks)
    o.set_frame(0, head_avatar.frames[0])
    rng = np.random.default_rng(11)
    pts = (0.5 + rng.uniform(-0.16, 0.16, (4000, 3))).astype(np.float32)
    ids, dist = o.closest(0, pts)
    bid, bd2 = o.closest_brute(0, pts)
    assert np.array_equal(ids, bid)
    assert np.array_equal(dist, np.sqrt(bd2).astype(np.float32))


def test_bvh_invariants(small_oracle, small_avatar):
    nodes, order = small_oracle.export_bvh(0)
    nodes = nodes.reshape(-1, 8)
    nf = len(small_avatar.faces)
    assert np.array_equal(np.sort(order), np.arange(nf))
    lo, hi = nodes[:, 0:3], nodes[:, 4:7]
    begin = nodes[:, 3].view(np.int32)
    end = nodes[:, 7].view(np.int32)
    n_nodes = len(nodes)
    leaf = begin >= 0
    # leaves partition [0, nf) in order
    lb, le = begin[leaf], end[leaf]
    assert lb[0] == 0 and le[-1] == nf and np.array_equal(lb[1:], le[:-1])
    s = small_oracle.cfg.bvh_leaf_min
    assert ((le - lb) >= s).all() and ((le - lb) <= 4 * s - 1).all()
    # inner nodes enclose their four children
    for i in np.nonzero(~leaf)[0]:
        ch = 4 * i + 1 + np.arange(4)
        assert (ch < n_nodes).all()
        assert (lo[i] <= lo[ch]).all() and (hi[i] >= hi[ch]).all()
    # leaves enclose their triangles
    verts = small_avatar.frames[0][small_avatar.faces[order]]
    for i in np.nonzero(leaf)[0][::7]:
        v = verts[begin[i]:end[i]].reshape(-1, 3)
        assert (lo[i] <= v.min(axis=0)).all() and (hi[i] >= v.max(axis=0)).all()


def test_adjacency_shares_edges(small_oracle, small_avatar, head_avatar):
    adj = small_oracle.adjacency()
    f = small_avatar.faces
    # closed sphere: exactly three edge neighbours, symmetric
    assert (adj >= 0).all()
    for i in range(0, len(f), 13):
        for j in adj[i]:
            assert len(set(f[i]) & set(f[j])) == 2
            assert i in adj[j]
        assert list(adj[i]) == sorted(adj[i])
    from oracle.oracle import Oracle
    from insta_amd.capi import full_config
    o = Oracle(full_config(max_frames=1))
    o.set_canonical_mesh(head_avatar.canonical, head_avatar.faces, head_avatar.masks)
    a2 = o.adjacency()
    # open neck ring + mouth slit: boundary triangles have fewer neighbours
    assert (a2 < 0).any() and ((a2 >= 0).sum(axis=1) >= 1).all()


def test_frames_orthonormal_and_inverse(small_oracle, small_avatar):
    P, Pi, cen = small_oracle.frames(-1)
    A = P[:, :, :3].astype(np.float64)
    eye = np.eye(3)
    assert np.abs(A.transpose(0, 2, 1) @ A - eye).max() < 1e-6
    assert np.abs(np.linalg.det(A) - 1.0).max() < 1e-6
    # translation is vertex a; column 2 is the unit normal
    tri = small_avatar.canonical[small_avatar.faces].astype(np.float64)
    assert np.abs(P[:, :, 3] - tri[:, 0]).max() == 0.0
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    assert np.abs(A[:, :, 2] - nrm).max() < 1e-6
    # inverse: [A|t] o [Ai|ti] = identity
    M = np.concatenate([P, np.tile([0, 0, 0, 1.0], (len(P), 1, 1))], axis=1)
    Mi = np.concatenate([Pi, np.tile([0, 0, 0, 1.0], (len(P), 1, 1))], axis=1)
    assert np.abs(M @ Mi - np.eye(4)).max() < 2e-6
    assert np.abs(cen - tri.mean(axis=1)).max() < 1e-6


def test_frames_eye_region_scaled_by_area_ratio(small_oracle, small_avatar):
    P, _, _ = small_oracle.frames(0)
    eye = small_avatar.masks == synth.MASK_EYE_REGION
    assert eye.any()
    tri = small_avatar.frames[0][small_avatar.faces].astype(np.float64)
    ctri = small_avatar.canonical[small_avatar.faces].astype(np.float64)
    area = np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
    carea = np.linalg.norm(np.cross(ctri[:, 1] - ctri[:, 0], ctri[:, 2] - ctri[:, 0]), axis=1)
    scale = np.linalg.norm(P[:, :, 0].astype(np.float64), axis=1)
    assert np.abs(scale[eye] - area[eye] / carea[eye]).max() < 1e-5
    assert np.abs(scale[~eye] - 1.0).max() < 1e-6


def _coords(pos, dirs_unit, frame, stride=11):
    c = np.zeros((len(pos), stride), np.float32)
    c[:, 0:3] = pos
    c[:, 4] = frame
    c[:, 5:8] = pos
    c[:, 8:11] = (dirs_unit + 1.0) * 0.5
    return c


def _near_surface(verts, faces, n, rng, spread=0.02):
    tri = verts[faces]
    k = rng.integers(0, len(faces), n)
    bary = rng.dirichlet([1, 1, 1], n)
    p = np.einsum("nk,nkd->nd", bary, tri[k])
    return (p + spread * rng.standard_normal((n, 3))).astype(np.float32)


def test_warp_identity_pose(small_avatar):
    from oracle.oracle import Oracle
    o = Oracle(small_config())
    o.set_canonical_mesh(small_avatar.canonical, small_avatar.faces, small_avatar.masks)
    o.set_frame(0, small_avatar.canonical)
    rng = np.random.default_rng(5)
    pos = _near_surface(small_avatar.canonical, small_avatar.faces, 4000, rng)
    d = rng.standard_normal((4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = _coords(pos, d, 0)
    out = o.deform(c)
    assert np.abs(out[:, 0:3] - pos).max() < 2e-6
    assert np.abs(out[:, 8:11] - c[:, 8:11]).max() < 2e-6


def test_warp_rigid_pose_is_rigid_inverse(small_avatar):
    from oracle.oracle import Oracle
    rng = np.random.default_rng(9)
    R = synth.random_rotation(rng)
    t = np.array([0.01, -0.02, 0.015])
    canon = small_avatar.canonical.astype(np.float64)
    posed = ((canon - 0.5) @ R.T + 0.5 + t).astype(np.float32)
    o = Oracle(small_config(disable_interp=1))
    o.set_canonical_mesh(small_avatar.canonical, small_avatar.faces, small_avatar.masks)
    o.set_frame(0, posed)
    pos = _near_surface(posed, small_avatar.faces, 4000, rng)
    d = rng.standard_normal((4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out = o.deform(_coords(pos, d, 0))
    expect = (pos.astype(np.float64) - 0.5 - t) @ R + 0.5
    assert np.abs(out[:, 0:3] - expect).max() < 5e-6
    dexp = d @ R
    assert np.abs(out[:, 8:11] * 2 - 1 - dexp).max() < 5e-6


def test_warp_holder_and_bounds(small_oracle, small_avatar):
    rng = np.random.default_rng(4)
    pos = _near_surface(small_avatar.frames[0], small_avatar.faces, 8, rng)
    d = np.tile([0, 0, 1.0], (8, 1))
    c = _coords(pos, d, 0)
    c[0, 4] = -1.0          # negative holder: untouched
    c[1, 4] = 7.0           # frame beyond the loaded frames: untouched
    c[2, 0] = 1.5           # outside the unit cube: untouched
    out = small_oracle.deform(c)
    assert np.array_equal(out[:3], c[:3])
    assert not np.array_equal(out[3:, :3], c[3:, :3])


# ---------------------------------------------------------------- sampling
def test_sobol_and_scramble_kats():
    for i in [0, 1, 2, 3, 5, 0x12345, 0xdeadbeef]:
        assert O.sobol(i, 0) == int(f"{i:032b}"[::-1], 2)
    # dimension 1: direction numbers 0x80000000, 0xc0000000, 0xa0000000, 0xf0000000, 0x88000000
    assert [O.sobol(1 << b, 1) for b in range(5)] == [0x80000000, 0xC0000000, 0xA0000000, 0xF0000000, 0x88000000]
    assert O.sobol(3, 1) == 0x80000000 ^ 0xC0000000
    # a nested uniform scramble keeps the first 2^m points stratified
    for seed in (0, 786433, 0xDEADBEEF):
        for m in (4, 8):
            v = np.array([O.ld_random_val(i, seed, 0) for i in range(1 << m)])
            assert np.array_equal(np.sort(np.floor(v * (1 << m)).astype(int)), np.arange(1 << m))
            assert (v >= 0).all() and (v < 1).all()
    off = O.pixel_offset(0)
    assert np.abs(off - 0.5).max() < 1e-7


def test_morton_round_trip():
    assert O.morton3(1, 0, 0) == 1 and O.morton3(0, 1, 0) == 2 and O.morton3(0, 0, 1) == 4
    assert O.morton3(1023, 1023, 1023) == (1 << 30) - 1
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1024, (200, 3)):
        m = O.morton3(int(x), int(y), int(z))
        assert (O.morton3_invert(m), O.morton3_invert(m >> 1), O.morton3_invert(m >> 2)) == (x, y, z)


def test_occupancy_matches_brute_force(small_avatar):
    from oracle.oracle import Oracle
    cfg = small_config(grid_size=16)
    o = Oracle(cfg)
    o.set_canonical_mesh(small_avatar.canonical, small_avatar.faces, small_avatar.masks)
    o.set_frame(0, small_avatar.frames[0])
    thr = 0.05
    bits = np.unpackbits(o.build_occupancy(0, thr), bitorder="little")
    G = 16
    idx = np.arange(G ** 3)
    xyz = np.stack([[O.morton3_invert(int(i) >> s) for s in (0, 1, 2)] for i in idx])
    centres = ((xyz + 0.5) / G).astype(np.float32)
    ids, d2 = o.closest_brute(0, centres)
    dist = np.sqrt(d2.astype(np.float64))
    expect = (dist < thr) & (small_avatar.masks[ids] != 0)
    clear = np.abs(dist - thr) > 1e-6
    assert np.array_equal(bits[clear].astype(bool), expect[clear])
    assert bits.sum() > 0


# ---------------------------------------------------------------- encoding
def test_encode_dense_levels_reproduce_linear_field():
    from oracle.oracle import Oracle
    cfg = small_config(n_levels=2, log2_hashmap_size=16)
    o = Oracle(cfg)
    lay = o.layout
    assert lay.level_dense[0] == 1 and lay.level_dense[1] == 1
    params = np.zeros(lay.n_params, np.float32)
    table = params[lay.n_mlp_params:].reshape(-1, 2)
    coef = np.array([0.3, -0.2, 0.1])
    for l in range(2):
        r = lay.level_resolution[l]
        g = np.arange(r)
        zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
        lin = (xx * coef[0] + yy * coef[1] + zz * coef[2]).ravel()
        idx = (xx + yy * r + zz * r * r).ravel()
        table[lay.level_offset[l] + idx, 0] = lin
        table[lay.level_offset[l] + idx, 1] = 1.0
    o.set_params(params)
    rng = np.random.default_rng(2)
    # with x * s + 0.5 the +1 corner wraps past the last vertex when
    # x > (r - 1.5) / s; stay below that on both levels
    x = rng.uniform(0.0, 0.95, (2000, 3)).astype(np.float32)
    f = o.encode(x)
    for l in range(2):
        s = lay.level_scale[l]
        pos = x.astype(np.float64) * s + 0.5
        assert np.abs(f[:, 2 * l] - (pos - 0.0) @ coef).max() < 1e-4
        # trilinear weights sum to one
        assert np.abs(f[:, 2 * l + 1] - 1.0).max() < 1e-6


def test_encode_hash_index_kats():
    from oracle.oracle import Oracle
    cfg = small_config()
    o = Oracle(cfg)
    lay = o.layout
    l = 7
    assert lay.level_dense[l] == 0
    T = lay.level_size[l]
    s = np.float32(lay.level_scale[l])
    # a position whose scaled coordinate floors to 0 on every axis
    x = np.float32(0.1) / s
    idx = o.encode_indices(np.array([[x, x, x]], np.float32))[0, l]
    assert idx[0] == 0
    assert idx[1] == 1 % T
    assert idx[2] == 2654435761 % T
    assert idx[4] == 805459861 % T
    assert idx[7] == (1 ^ 2654435761 ^ 805459861) % T
    # level scales: b^l * N_min - 1, resolution ceil(scale) + 1
    for l in range(cfg.n_levels):
        assert math.isclose(lay.level_scale[l], 1.5 ** l * 16 - 1, rel_tol=1e-6)
        assert lay.level_resolution[l] == math.ceil(lay.level_scale[l]) + 1


def test_encode_backward_is_adjoint_of_forward(small_oracle):
    rng = np.random.default_rng(8)
    x = rng.uniform(0.05, 0.95, (300, 3)).astype(np.float32)
    dL = rng.standard_normal((300, small_oracle.cfg.n_levels * 2)).astype(np.float32)
    g = small_oracle.encode_backward(x, dL)
    # <dL, enc(table)> is linear in the table: its gradient is g
    lay = small_oracle.layout
    base = small_oracle
    params = synth.network_params(7, lay, base.cfg)
    e0 = base.encode(x).astype(np.float64)
    t = rng.standard_normal(lay.n_grid_params).astype(np.float32) * 1e-3
    p2 = params.copy()
    p2[lay.n_mlp_params:] += t
    base.set_params(p2)
    e1 = base.encode(x).astype(np.float64)
    base.set_params(params)
    lhs = np.sum(dL * (e1 - e0))
    rhs = float(g @ t.astype(np.float64))
    assert math.isclose(lhs, rhs, rel_tol=1e-3)


# ---------------------------------------------------------------- network
def test_sh_orthonormal():
    n = 200000
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    theta = math.pi * (1 + 5 ** 0.5) * i
    d = np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], axis=1)
    Y = O.sh16(((d + 1) * 0.5).astype(np.float32)).astype(np.float64)
    G = Y.T @ Y * (4 * math.pi / n)
    assert np.abs(G - np.eye(16)).max() < 2e-3


def test_network_composition_matches_matrix_form(small_oracle):
    cfg = small_oracle.cfg
    lay = small_oracle.layout
    params = synth.network_params(7, lay, cfg).astype(np.float64)
    rng = np.random.default_rng(12)
    n = 500
    pos = rng.uniform(0.1, 0.9, (n, 3)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = _coords(pos, d, 0)
    out = small_oracle.network(c)
    enc = small_oracle.encode(pos).astype(np.float64)
    sh = O.sh16(c[:, 8:11]).astype(np.float64)
    off = 0
    mats = []
    for (fo, fi) in lay.mlp_shapes(cfg):
        mats.append(params[off:off + fo * fi].reshape(fo, fi))
        off += fo * fi
    relu = lambda v: np.maximum(v, 0)
    h = relu(enc @ mats[0].T)
    dout = h @ mats[1].T
    r_in = np.concatenate([dout, sh], axis=1)
    h = relu(r_in @ mats[2].T)
    h = relu(h @ mats[3].T)
    rgb = h @ mats[4].T
    ref = np.concatenate([rgb, dout[:, :1]], axis=1)
    assert_rel_close(out, ref, 2e-5, "network")


# ---------------------------------------------------------------- march + composite
def _render_scene(min_t=0.0, zero_mlp=False, full=False, empty=False):
    from oracle.oracle import Oracle
    av = synth.small_sphere_avatar(0, 1)
    cfg = small_config(min_transmittance=min_t)
    o = Oracle(cfg)
    o.set_canonical_mesh(av.canonical, av.faces, av.masks)
    o.set_frame(0, av.frames[0])
    p = synth.network_params(3, o.layout, cfg)
    if zero_mlp:
        p[:o.layout.n_mlp_params] = 0.0
    o.set_params(p)
    G = cfg.grid_size
    if full or empty:
        o.set_occupancy(0, np.full(G ** 3 // 8, 0xFF if full else 0, np.uint8))
    else:
        o.build_occupancy(0, 0.05)
    cam = make_camera(synth.look_at_camera(np.array([0.2, 0.3, 1.0]), 1.0), synth.focal_from_fov(32, 30.0))
    return o, cfg, cam


def test_march_empty_grid_gives_nothing():
    o, cfg, cam = _render_scene(empty=True)
    rgba, depth, steps, stats = o.render(cam, 32, 32, 0, 0)
    assert steps.sum() == 0 and (rgba == 0).all() and (depth == 1e10).all()


def test_march_full_grid_step_count_closed_form():
    o, cfg, cam = _render_scene(full=True, zero_mlp=True)
    rgba, depth, steps, stats = o.render(cam, 32, 32, 0, 0)
    box = o.render_box().astype(np.float64)
    lo, hi = box[:3], box[3:]
    M = np.array(cam.matrix, np.float64).reshape(3, 4)
    hit = 0
    for i in range(32 * 32):
        x, y = i % 32, i // 32
        u, v = (x + 0.5) / 32, (y + 0.5) / 32
        d = M[:, :3] @ np.array([(u - 0.5) * 32 / cam.focal_px[0], (v - 0.5) * 32 / cam.focal_px[1], 1.0])
        d /= np.linalg.norm(d)
        o_ = M[:, 3]
        with np.errstate(divide="ignore"):
            t1, t2 = (lo - o_) / d, (hi - o_) / d
        tmin, tmax = np.minimum(t1, t2).max(), np.maximum(t1, t2).min()
        if tmin >= tmax:
            assert steps[i] == 0
            continue
        hit += 1
        t0 = max(tmin, 0.05) + 1e-6
        jitter = O.ld_random_val(0, i * 786433, 0) * cfg.dt
        expect = min(math.floor((tmax - t0 - jitter) / cfg.dt) + 1, cfg.max_steps)
        assert abs(int(steps[i]) - expect) <= 1, (i, steps[i], expect)
    assert hit > 100


def test_composite_constant_density_closed_form():
    o, cfg, cam = _render_scene(full=True, zero_mlp=True)
    rgba, depth, steps, stats = o.render(cam, 32, 32, 0, 0)
    sigma = math.exp(cfg.density_logit_bias)
    alpha = 1 - math.exp(-sigma * cfg.dt)
    n = steps.astype(np.float64)
    w_expect = 1 - (1 - alpha) ** n
    # pixels with alpha <= 0.001 are cleared by the compaction rule
    keep = w_expect > 0.0011
    assert np.abs(rgba[keep, 3] - w_expect[keep]).max() < 2e-5
    assert np.abs(rgba[keep, 0] - 0.5 * rgba[keep, 3]).max() < 1e-6
    # first sample carries the largest weight: depth is the first sample's
    assert (depth[n == 0] == 1e10).all()


def test_composite_early_termination_normalises():
    o, cfg, cam = _render_scene(min_t=0.5, full=True, zero_mlp=True)
    rgba, depth, steps, stats = o.render(cam, 32, 32, 0, 0)
    sigma = math.exp(cfg.density_logit_bias)
    alpha = 1 - math.exp(-sigma * cfg.dt)
    n_stop = math.ceil(math.log(0.5) / math.log(1 - alpha) - 1e-9)
    term = rgba[:, 3] > 0.999
    assert term.any()
    assert (np.abs(steps[term].astype(int) - n_stop) <= 1).all()
    assert np.abs(rgba[term, 3] - 1.0).max() < 1e-6
