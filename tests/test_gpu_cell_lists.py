"""Nearest-triangle candidate lists (celllist.hip) against brute force.

A frame whose occupancy was built with insta_build_occupancy answers queries
inside its occupied cells from per-sub-cell candidate lists.  The answer must
be the brute-force argmin under the (squared distance, face id) order,
exactly as the BVH's: ids and distances bit-exact against the oracle's
brute-force search, for points spread over occupied cells, on sub-cell and
cell faces and corners (where a point's sub-cell is decided by rounding),
and for points outside (which take the BVH).  Renders with and without lists
are identical, lists are dropped when the frame's mesh changes, and
capacities too small for a frame fall back to the BVH without changing
results."""
import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import full_config, make_camera, small_config
from tests.conftest import ensure_oracle

pytestmark = pytest.mark.gpu


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _compact3(v):
    v = v & 0x09249249
    v = (v | (v >> 2)) & 0x030C30C3
    v = (v | (v >> 4)) & 0x0300F00F
    v = (v | (v >> 8)) & 0x030000FF
    v = (v | (v >> 16)) & 0x000003FF
    return v


def _occupied_cells(bits):
    """(x, y, z) of the set bits of a Morton-order bitfield."""
    idx = np.nonzero(np.unpackbits(np.asarray(bits, np.uint8), bitorder="little"))[0].astype(np.uint32)
    return np.stack([_compact3(idx), _compact3(idx >> 1), _compact3(idx >> 2)], 1)


def _in_occupied(pts, cells, G):
    """Points whose cell, by the march's float32 lookup, is one of `cells`."""
    f32 = np.float32
    g = ((pts.astype(f32) - f32(0.5)) * f32(1.0) + f32(0.5)) * f32(G)
    c = np.minimum(g.astype(np.int64), G - 1)
    key = c[:, 0] + G * (c[:, 1] + G * c[:, 2])
    occ = cells[:, 0].astype(np.int64) + G * (cells[:, 1] + G * cells[:, 2].astype(np.int64))
    return np.isin(key, occ)


def _scene(cfg, av, threshold, n_frames=1):
    torch = _torch()
    ensure_oracle()
    from insta_amd.capi import Context
    from oracle.oracle import Oracle
    ctx, orc = Context(cfg), Oracle(cfg)
    for s in (ctx, orc):
        s.set_canonical_mesh(av.canonical, av.faces, av.masks)
    ctx.set_frames(0, torch.from_numpy(np.stack(av.frames[:n_frames])).cuda())
    for f in range(n_frames):
        orc.set_frame(f, av.frames[f])
    bits = [orc.build_occupancy(f, threshold) for f in range(n_frames)]
    for f in range(n_frames):
        ctx.build_occupancy(f, threshold)
    torch.cuda.synchronize()
    return ctx, orc, bits


def _probe_points(cells, G, S, rng, n_random):
    """Random points in occupied cells, sub-cell lattice points (faces, edges
    and corners of sub-cells and cells) of a sample of occupied cells, and
    points nudged one float step off them."""
    c = cells[rng.integers(0, len(cells), n_random)]
    pts = [(c + rng.uniform(0.0, 1.0, (n_random, 3))) / G]
    pick = cells[rng.choice(len(cells), min(len(cells), 400), replace=False)]
    lat = np.stack(np.meshgrid(*[np.arange(S + 1) / S] * 3, indexing="ij"), -1).reshape(-1, 3)
    on = ((pick[:, None, :] + lat[None]) / G).reshape(-1, 3)
    pts += [on, np.nextafter(on.astype(np.float32), np.float32(2)), np.nextafter(on.astype(np.float32), np.float32(-1))]
    # mixed: one coordinate on a face, the others random
    k = rng.integers(0, 3, n_random)
    m = (c + rng.uniform(0.0, 1.0, (n_random, 3))) / G
    m[np.arange(n_random), k] = (c[np.arange(n_random), k] + rng.integers(0, S + 1, n_random) / S) / G
    pts.append(m)
    return np.clip(np.concatenate(pts), 0.0, 1.0).astype(np.float32)


# list capacities large enough that every sub-cell gets its list
CASES = {
    "small": (lambda S: small_config(max_frames=1, cell_split=S, cell_list_entries=1 << 25),
              lambda: synth.small_sphere_avatar(seed=0, n_frames=1), 0.05),
    "head": (lambda S: full_config(max_frames=1, cell_split=S, cell_list_entries=1 << 25),
             lambda: synth.head_avatar(seed=1, n_frames=1), 0.015),
}


@pytest.mark.parametrize("split", [1, 2, 4])
@pytest.mark.parametrize("which", ["small", "head"])
def test_list_ids_equal_brute_force(which, split):
    torch = _torch()
    make_cfg, make_av, thr = CASES[which]
    cfg = make_cfg(split)
    ctx, orc, bits = _scene(cfg, make_av(), thr)
    cells = _occupied_cells(bits[0])
    assert len(cells) > 100
    rng = np.random.default_rng(split * 10 + len(which))
    pts = _probe_points(cells, cfg.grid_size, split, rng, 40000)
    pts = np.concatenate([pts, rng.uniform(0.0, 1.0, (20000, 3)).astype(np.float32)])  # mostly unoccupied: BVH
    info = ctx.cell_list_info(0)
    n_sub = len(cells) * split ** 3
    assert info["valid"] == 1 and info["occupied_cells"] == len(cells), info
    # a list longer than the sort's reach leaves its sub-cell to the BVH
    assert info["subcells_without_list"] <= n_sub // 20, info
    ctx.set_profiling(True, count_traversal=True)
    ids, dist = ctx.closest(0, torch.from_numpy(pts).cuda())
    torch.cuda.synchronize()
    tc = ctx.traversal_counts()
    ctx.set_profiling(False)
    # the points whose cell (the march's lookup, in float32) is occupied take
    # a list, unless their sub-cell has none
    expect = int(_in_occupied(pts, cells, cfg.grid_size).sum())
    if info["subcells_without_list"] == 0:
        assert tc["list_queries"] == expect, tc
    else:
        assert 0.8 * expect <= tc["list_queries"] < expect, (tc, expect)
    rid, rd2 = orc.closest_brute(0, pts)
    ids = ids.cpu().numpy()
    bad = np.nonzero(ids != rid)[0]
    assert len(bad) == 0, f"{len(bad)} ids differ, first at {pts[bad[0]]}: {ids[bad[0]]} vs {rid[bad[0]]}"
    assert np.array_equal(dist.cpu().numpy(), np.sqrt(rd2))


def test_render_with_lists_equals_bvh_only_render(head_avatar):
    """Every split renders the same image, bit for bit, as the BVH alone; the
    counting build shows the samples took the lists."""
    torch = _torch()
    from insta_amd.capi import Context
    cam = make_camera(synth.random_camera(np.random.default_rng(3), 0.7), synth.focal_from_fov(256, 30.0))
    out = {}
    for split in (0, 1, 2, 4):
        cfg = full_config(max_frames=2, cell_split=split, cell_list_entries=1 << 25)
        ctx = Context(cfg)
        ctx.set_canonical_mesh(head_avatar.canonical, head_avatar.faces, head_avatar.masks)
        ctx.set_frames(0, torch.from_numpy(np.stack(head_avatar.frames[:2])).cuda())
        ctx.set_params(torch.from_numpy(synth.network_params(7, ctx.layout, cfg)).cuda())
        ctx.build_occupancy(0, 0.015)
        info = ctx.cell_list_info(0)
        assert info["valid"] == (1 if split else 0), (split, info)
        ctx.set_profiling(True, count_traversal=True)
        rgba, depth, steps = ctx.render(cam, 256, 256, 0, 0)
        torch.cuda.synchronize()
        tc = ctx.traversal_counts()
        ctx.set_profiling(False)
        out[split] = (rgba.cpu().numpy(), depth.cpu().numpy(), steps.cpu().numpy(), tc)
        assert tc["queries"] > 100000
        if split == 0:
            assert tc["list_queries"] == 0
        elif info["subcells_without_list"] == 0:
            # every marched sample lies in an occupied cell of its frame's grid
            assert tc["list_queries"] == tc["queries"], tc
        else:
            assert 0.8 * tc["queries"] <= tc["list_queries"] < tc["queries"], (tc, info)
        del ctx
    for split in (1, 2, 4):
        for a, b in zip(out[0][:3], out[split][:3]):
            assert np.array_equal(a, b), split


def test_lists_dropped_when_the_mesh_changes(head_avatar):
    """A frame rebuilt with a new expression answers from its new mesh even
    before its occupancy (and lists) are rebuilt."""
    torch = _torch()
    cfg = full_config(max_frames=2, cell_split=2)
    ctx, orc, bits = _scene(cfg, head_avatar, 0.015)
    rng = np.random.default_rng(8)
    coeffs = (3.0 * rng.standard_normal(head_avatar.bases.shape[0])).astype(np.float32)
    ctx.set_frame_expression(0, torch.from_numpy(head_avatar.bases).cuda(), torch.from_numpy(coeffs).cuda())
    v = head_avatar.canonical.astype(np.float32).copy()
    for k in range(len(coeffs)):
        v = (v + coeffs[k] * head_avatar.bases[k]).astype(np.float32)
    orc.set_frame(0, v)
    pts = _probe_points(_occupied_cells(bits[0]), cfg.grid_size, 2, rng, 30000)
    ids, _ = ctx.closest(0, torch.from_numpy(pts).cuda())
    rid, _ = orc.closest_brute(0, pts)
    assert np.array_equal(ids.cpu().numpy(), rid)
    # and once rebuilt, the lists of the new mesh give the same answers
    ctx.build_occupancy(0, 0.015)
    ids2, _ = ctx.closest(0, torch.from_numpy(pts).cuda())
    assert np.array_equal(ids2.cpu().numpy(), rid)


@pytest.mark.parametrize("caps", [dict(cell_list_entries=5000), dict(cell_list_cells=50)])
def test_small_capacities_fall_back_to_the_bvh(head_avatar, caps):
    torch = _torch()
    cfg = full_config(max_frames=2, cell_split=2, **caps)
    ctx, orc, bits = _scene(cfg, head_avatar, 0.015)
    rng = np.random.default_rng(9)
    pts = _probe_points(_occupied_cells(bits[0]), cfg.grid_size, 2, rng, 30000)
    ids, dist = ctx.closest(0, torch.from_numpy(pts).cuda())
    rid, rd2 = orc.closest_brute(0, pts)
    assert np.array_equal(ids.cpu().numpy(), rid)
    assert np.array_equal(dist.cpu().numpy(), np.sqrt(rd2))
