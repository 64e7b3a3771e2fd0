"""GPU parity: the HIP library through the C ABI against the CPU oracle on
identical seeded inputs.

Bar: nearest-triangle ids, BVH layout, occupancy bits, hash-grid features and
march step counts are bit-exact; FP32 positions, directions, network outputs,
composited pixels and depths are within 1e-5 relative (relative to
|ref| + mean|ref|, so entries near zero are judged on the array's scale).
"""
import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import full_config, make_camera, small_config
from tests.conftest import assert_rel_close, assert_render_matches, ensure_oracle

pytestmark = pytest.mark.gpu
RTOL = 1e-5


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _scene(cfg, avatar, params_seed=7):
    """The same avatar in the HIP library and in the oracle."""
    torch = _torch()
    ensure_oracle()
    from insta_amd.capi import Context
    from oracle.oracle import Oracle
    ctx = Context(cfg)
    orc = Oracle(cfg)
    ctx.set_canonical_mesh(avatar.canonical, avatar.faces, avatar.masks)
    orc.set_canonical_mesh(avatar.canonical, avatar.faces, avatar.masks)
    frames = np.stack(avatar.frames).astype(np.float32)
    ctx.set_frames(0, torch.from_numpy(frames).cuda())
    for f, v in enumerate(avatar.frames):
        orc.set_frame(f, v)
    params = synth.network_params(params_seed, ctx.layout, cfg)
    ctx.set_params(torch.from_numpy(params).cuda())
    orc.set_params(params)
    torch.cuda.synchronize()
    return ctx, orc


@pytest.fixture(scope="module")
def small(small_avatar):
    return _scene(small_config(), small_avatar)


@pytest.fixture(scope="module")
def head(head_avatar):
    return _scene(full_config(max_frames=4), head_avatar)


def _points(rng, n, lo=0.0, hi=1.0):
    return rng.uniform(lo, hi, (n, 3)).astype(np.float32)


def _near_surface(verts, faces, n, rng, spread=0.02):
    tri = verts[faces].astype(np.float64)
    k = rng.integers(0, len(faces), n)
    bary = rng.dirichlet([1, 1, 1], n)
    p = np.einsum("nk,nkd->nd", bary, tri[k])
    return (p + spread * rng.standard_normal((n, 3))).astype(np.float32)


def _coords(pos, dirs_unit, frame, stride=11):
    c = np.zeros((len(pos), stride), np.float32)
    c[:, 0:3] = pos
    c[:, 4] = frame
    c[:, 5:8] = pos
    c[:, 8:11] = (dirs_unit + 1.0) * 0.5
    return c


# ---------------------------------------------------------------- BVH
@pytest.mark.parametrize("frame", [-1, 0, 1])
def test_closest_ids_bit_exact_small(small, small_avatar, frame):
    torch = _torch()
    ctx, orc = small
    rng = np.random.default_rng(100 + frame)
    pts = np.concatenate([_points(rng, 60000), _near_surface(small_avatar.canonical, small_avatar.faces, 60000, rng)])
    ids, dist = ctx.closest(frame, torch.from_numpy(pts).cuda())
    rid, rdist = orc.closest(frame, pts)
    assert np.array_equal(ids.cpu().numpy(), rid)
    assert np.array_equal(dist.cpu().numpy(), rdist)


def test_closest_ids_bit_exact_head(head, head_avatar):
    torch = _torch()
    ctx, orc = head
    rng = np.random.default_rng(5)
    pts = np.concatenate([_points(rng, 50000, 0.3, 0.7),
                          _near_surface(head_avatar.frames[0], head_avatar.faces, 50000, rng, 0.01)])
    for frame in (-1, 0, 3):
        ids, dist = ctx.closest(frame, torch.from_numpy(pts).cuda())
        rid, rdist = orc.closest(frame, pts)
        assert np.array_equal(ids.cpu().numpy(), rid), frame
        assert np.array_equal(dist.cpu().numpy(), rdist), frame


@pytest.mark.parametrize("which", ["small", "head"])
def test_device_bvh_equals_oracle_bvh(which, request):
    ctx, orc = request.getfixturevalue(which)
    for frame in (-1, 0):
        nodes, order = ctx.export_bvh(frame)
        rn, ro = orc.export_bvh(frame)
        assert np.array_equal(order.cpu().numpy().view(np.uint32), ro)
        assert np.array_equal(nodes.cpu().numpy().view(np.uint32), rn.view(np.uint32))


def test_expression_rebuild_matches_host_blend(head, head_avatar):
    torch = _torch()
    ctx, orc = head
    rng = np.random.default_rng(21)
    coeffs = rng.standard_normal(head_avatar.bases.shape[0]).astype(np.float32)
    ctx.set_frame_expression(2, torch.from_numpy(head_avatar.bases).cuda(), torch.from_numpy(coeffs).cuda())
    # the same float32 accumulation order on the host
    v = head_avatar.canonical.astype(np.float32).copy()
    for k in range(len(coeffs)):
        v = (v + coeffs[k] * head_avatar.bases[k]).astype(np.float32)
    orc.set_frame(2, v)
    pts = _near_surface(v, head_avatar.faces, 40000, rng, 0.01)
    ids, dist = ctx.closest(2, torch.from_numpy(pts).cuda())
    rid, rdist = orc.closest(2, pts)
    assert np.array_equal(ids.cpu().numpy(), rid)
    assert np.array_equal(dist.cpu().numpy(), rdist)
    # restore frame 2 for later tests
    ctx.set_frames(2, torch.from_numpy(head_avatar.frames[2][None]).cuda())
    orc.set_frame(2, head_avatar.frames[2])


# ---------------------------------------------------------------- deform
@pytest.mark.parametrize("which", ["small", "head"])
def test_deform_matches_oracle(which, request):
    torch = _torch()
    ctx, orc = request.getfixturevalue(which)
    av = request.getfixturevalue("small_avatar" if which == "small" else "head_avatar")
    rng = np.random.default_rng(31)
    n = 80000
    nfr = len(av.frames)
    frame = rng.integers(0, nfr, n)
    pos = np.concatenate([_near_surface(av.frames[f], av.faces, n // nfr + 1, rng, 0.015) for f in range(nfr)])[:n]
    frame = np.repeat(np.arange(nfr), n // nfr + 1)[:n]
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = _coords(pos, d, 0)
    c[:, 4] = frame
    c[::97, 4] = -1.0                    # skipped
    c[::89, 0] = 1.25                    # outside the unit cube: skipped
    g = torch.from_numpy(c).cuda()
    ctx.deform_samples(g, -1)
    ref = orc.deform(c, -1)
    got = g.cpu().numpy()
    skipped = (c[:, 4] < 0) | (c[:, 0] > 1)
    assert np.array_equal(got[skipped], c[skipped])
    assert_rel_close(got[:, 0:3], ref[:, 0:3], RTOL, "canonical position")
    assert_rel_close(got[:, 8:11], ref[:, 8:11], RTOL, "canonical direction")
    assert np.array_equal(got[:, [3, 4, 5, 6, 7]], ref[:, [3, 4, 5, 6, 7]])
    # fixed-frame mode ignores the holder's frame id
    g2 = torch.from_numpy(c).cuda()
    ctx.deform_samples(g2, 0)
    assert_rel_close(g2.cpu().numpy()[:, 0:3], orc.deform(c, 0)[:, 0:3], RTOL, "fixed frame")


# ---------------------------------------------------------------- occupancy
@pytest.mark.parametrize("which,thr", [("small", 0.05), ("head", 0.015)])
def test_occupancy_bit_exact(which, thr, request):
    ctx, orc = request.getfixturevalue(which)
    for frame in (-1, 0):
        bits = ctx.build_occupancy(frame, thr).cpu().numpy()
        ref = orc.build_occupancy(frame, thr)
        assert np.array_equal(bits, ref), frame
        assert np.unpackbits(ref).sum() > 1000


# ---------------------------------------------------------------- network
@pytest.mark.parametrize("which", ["small", "head"])
def test_encode_bit_exact(which, request):
    torch = _torch()
    ctx, orc = request.getfixturevalue(which)
    rng = np.random.default_rng(41)
    x = _points(rng, 100000, 0.02, 0.98)
    got = ctx.encode(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = orc.encode(x)
    assert np.array_equal(got, ref)


def test_encode_backward_matches_oracle(small):
    torch = _torch()
    ctx, orc = small
    rng = np.random.default_rng(43)
    x = _points(rng, 20000, 0.02, 0.98)
    dL = rng.standard_normal((20000, ctx.cfg.n_levels * ctx.cfg.n_features_per_level)).astype(np.float32)
    grad = torch.zeros(ctx.layout.n_grid_params, dtype=torch.float32, device="cuda")
    ctx.encode_backward(torch.from_numpy(x).cuda(), torch.from_numpy(dL).cuda(), grad)
    ref = orc.encode_backward(x, dL)
    got = grad.cpu().numpy().astype(np.float64)
    touched = ref != 0
    assert touched.sum() > 10000
    scale = np.abs(ref[touched]).mean()
    assert np.abs(got - ref).max() <= RTOL * (np.abs(ref).max() + scale)


@pytest.mark.parametrize("which", ["small", "head"])
def test_network_matches_oracle(which, request):
    torch = _torch()
    ctx, orc = request.getfixturevalue(which)
    rng = np.random.default_rng(51)
    n = 60000
    pos = _points(rng, n, 0.05, 0.95)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = _coords(pos, d, 0)
    got = ctx.network_inference(torch.from_numpy(c).cuda()).cpu().numpy()
    ref = orc.network(c)
    assert_rel_close(got, ref, RTOL, "network outputs")


# ---------------------------------------------------------------- render
def _render_pair(ctx, orc, cam, w, h, frame, occ, pixels=None, spp=0):
    rgba, depth, steps = ctx.render(cam, w, h, frame, occ, spp)
    rgba, depth, steps = rgba.cpu().numpy(), depth.cpu().numpy(), steps.cpu().numpy()
    r_rgba, r_depth, r_steps, stats = orc.render(cam, w, h, frame, occ, spp, pixels)
    if pixels is not None:
        rgba, depth, steps = rgba[pixels], depth[pixels], steps[pixels]
    return (rgba, depth, steps.astype(np.uint32)), (r_rgba, r_depth, r_steps), stats


def test_render_small_geometry_exact(small_avatar):
    """No early termination: per-pixel step counts are pure march geometry."""
    ctx, orc = _scene(small_config(min_transmittance=0.0), small_avatar)
    ctx.build_occupancy(0, 0.05)
    orc.build_occupancy(0, 0.05)
    cam = make_camera(synth.look_at_camera(np.array([0.3, 0.2, 1.0]), 1.0), synth.focal_from_fov(64, 30.0))
    (rgba, depth, steps), (r_rgba, r_depth, r_steps), stats = _render_pair(ctx, orc, cam, 64, 64, 0, 0)
    assert np.array_equal(steps, r_steps)
    assert steps.sum() == stats[0] and steps.sum() > 10000
    assert_rel_close(rgba, r_rgba, RTOL, "rgba")
    hit = r_depth < 1e9
    assert np.array_equal(depth >= 1e9, ~hit)
    assert_rel_close(depth[hit], r_depth[hit], RTOL, "depth")
    st = ctx.render_stats()
    assert st.n_samples == stats[0] and st.n_rays_alive0 == stats[1]


def test_render_small_default(small):
    ctx, orc = small
    ctx.build_occupancy(0, 0.05)
    orc.build_occupancy(0, 0.05)
    for k, spp in enumerate((0, 3)):
        cam = make_camera(synth.look_at_camera(np.array([0.1, -0.3, 1.0 - k]), 1.0), synth.focal_from_fov(64, 30.0),
                          snap=(spp == 0))
        (rgba, depth, steps), (r_rgba, r_depth, r_steps), _ = _render_pair(ctx, orc, cam, 64, 64, 0, 0, spp=spp)
        # a ray whose accumulated alpha lands within rounding of the
        # termination threshold may stop one sample apart (bounded there)
        assert_render_matches(rgba, steps, r_rgba, r_steps, RTOL, ctx.cfg.min_transmittance, "rgba")


def test_render_full_size_subset(head, head_avatar):
    """512x512 full-size render; a bounded subset of pixels is checked
    against the oracle, every pixel against the properties below."""
    ctx, orc = head
    ctx.build_occupancy(0, 0.015)
    orc.build_occupancy(0, 0.015)
    cam = make_camera(synth.look_at_camera(np.array([0.2, 0.1, 1.0]), 0.7), synth.focal_from_fov(512, 30.0))
    rng = np.random.default_rng(61)
    pixels = rng.choice(512 * 512, 3000, replace=False).astype(np.uint32)
    (rgba, depth, steps), (r_rgba, r_depth, r_steps), _ = _render_pair(ctx, orc, cam, 512, 512, 0, 0, pixels)
    assert np.array_equal(steps, r_steps)
    assert_rel_close(rgba, r_rgba, RTOL, "rgba")
    hit = r_depth < 1e9
    assert_rel_close(depth[hit], r_depth[hit], RTOL, "depth")
    full, _, fsteps = ctx.render(cam, 512, 512, 0, 0)
    full = full.cpu().numpy()
    st = ctx.render_stats()
    assert st.n_samples == int(fsteps.cpu().numpy().astype(np.int64).sum())
    assert (full >= 0).all() and (full[:, 3] <= 1.0 + 1e-6).all()
    assert (full[:, :3] <= full[:, 3:4] + 1e-6).all()


def test_render_is_deterministic(head):
    torch = _torch()
    ctx, orc = head
    ctx.build_occupancy(1, 0.015)
    cam = make_camera(synth.look_at_camera(np.array([-0.3, 0.2, 1.0]), 0.7), synth.focal_from_fov(256, 30.0))
    a = [t.clone() for t in ctx.render(cam, 256, 256, 1, 1)]
    b = ctx.render(cam, 256, 256, 1, 1)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_results_do_not_depend_on_the_bvh_leaf_size(head_avatar):
    """The nearest-triangle order (dist^2, original id) makes ids, occupancy
    and renders independent of the tree layout: leaves of 1..3 and 8..31
    triangles give identical outputs."""
    torch = _torch()
    from insta_amd.capi import Context
    rng = np.random.default_rng(77)
    pts = torch.from_numpy(_points(rng, 100000, 0.3, 0.7)).cuda()
    cam = make_camera(synth.look_at_camera(np.array([0.4, 0.1, 1.0]), 0.7), synth.focal_from_fov(128, 30.0))
    out = []
    for leaf in (1, 8):
        cfg = full_config(max_frames=2, bvh_leaf_min=leaf)
        ctx = Context(cfg)
        ctx.set_canonical_mesh(head_avatar.canonical, head_avatar.faces, head_avatar.masks)
        ctx.set_frames(0, torch.from_numpy(np.stack(head_avatar.frames[:2])).cuda())
        ctx.set_params(torch.from_numpy(synth.network_params(7, ctx.layout, cfg)).cuda())
        occ = ctx.build_occupancy(1, 0.015)
        ids, dist = ctx.closest(1, pts)
        rgba, _, steps = ctx.render(cam, 128, 128, 1, 1)
        torch.cuda.synchronize()
        out.append([t.cpu().numpy() for t in (occ, ids, dist, rgba, steps)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("split", [2, 0])
def test_traversal_counting_build_renders_the_same(head, head_avatar, split):
    """The counting build of the deformation kernel (profiling level 2) gives
    the same image, and its counts obey the query's structure.  BVH only
    (cell_split 0): one query per deformed sample, at least k inner nodes (the
    first descent) and one leaf per query, at most 15 exact triangle
    distances per leaf (a triangle whose own box is beyond the bound is
    skipped) and at least one per query.  With candidate lists: every query
    takes its sub-cell's list, no node is visited, and each list query scans
    at least its seed entry and measures at least that triangle."""
    torch = _torch()
    if split:
        ctx, _ = head
    else:
        ctx, _ = _scene(full_config(max_frames=4, cell_split=0), head_avatar)
    ctx.build_occupancy(1, 0.015)
    cam = make_camera(synth.look_at_camera(np.array([0.2, -0.1, 1.0]), 0.7), synth.focal_from_fov(192, 30.0))
    a = [t.clone() for t in ctx.render(cam, 192, 192, 1, 1)]
    ctx.set_profiling(True, count_traversal=True)
    b = ctx.render(cam, 192, 192, 1, 1)
    st = ctx.render_stats()
    tc = ctx.traversal_counts()
    ctx.set_profiling(False)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    assert 0 < tc["queries"] <= st.n_samples_emitted
    if split:
        assert tc["list_queries"] == tc["queries"]
        assert tc["inner"] == 0 and tc["leaves"] == 0
        assert tc["queries"] <= tc["triangles"] <= tc["list_entries"]
    else:
        k = int(np.floor(np.log(9976) / np.log(4))) - 1   # leaves of 4..15 triangles
        assert tc["list_queries"] == 0 and tc["list_entries"] == 0
        assert tc["inner"] >= k * tc["queries"]
        assert tc["leaves"] >= tc["queries"]
        assert tc["queries"] <= tc["triangles"] <= 15 * tc["leaves"]
    # counting is off again: counters are no longer reported
    assert ctx.traversal_counts()["queries"] == 0


# ---------------------------------------------------------------- other shapes and precisions
def test_network_fallback_shape_matches_oracle(small_avatar):
    """Two density hidden layers: not an MFMA shape, served by the scalar kernel."""
    torch = _torch()
    ctx, orc = _scene(small_config(density_hidden_layers=2), small_avatar)
    rng = np.random.default_rng(52)
    pos = _points(rng, 20000, 0.05, 0.95)
    d = rng.standard_normal((20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = _coords(pos, d, 0)
    got = ctx.network_inference(torch.from_numpy(c).cuda()).cpu().numpy()
    assert_rel_close(got, orc.network(c), RTOL, "network (scalar kernel)")


def _psnr(a, b):
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10 * np.log10(1.0 / mse)


@pytest.mark.parametrize("prec,tol,rms_tol", [(1, 2e-2, 2e-3), (2, 1.5e-1, 1.5e-2)])
def test_network_reduced_precision(head_avatar, prec, tol, rms_tol):
    """FP16 tables + FP16/BF16 MFMA MLPs with FP32 accumulation against the
    FP32 oracle.  Five chained layers each round their inputs (2^-11 FP16,
    2^-8 BF16); the bounds are on error / mean |output| per channel."""
    torch = _torch()
    ctx, orc = _scene(full_config(max_frames=4, precision=prec), head_avatar)
    rng = np.random.default_rng(53)
    n = 40000
    pos = _points(rng, n, 0.05, 0.95)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = _coords(pos, d, 0)
    got = ctx.network_inference(torch.from_numpy(c).cuda()).cpu().numpy().astype(np.float64)
    ref = orc.network(c).astype(np.float64)
    scale = np.abs(ref).mean(axis=0)
    rel = np.abs(got - ref) / scale
    err, rms = rel.max(), np.sqrt((rel ** 2).mean())
    print(f"precision {prec}: error / mean |ref|: max {err:.2e} rms {rms:.2e}")
    assert err < tol and rms < rms_tol


@pytest.mark.parametrize("prec", [1, 2])
def test_render_reduced_precision_psnr(head_avatar, prec):
    ctx, orc = _scene(full_config(max_frames=4, precision=prec), head_avatar)
    ctx.build_occupancy(0, 0.015)
    orc.build_occupancy(0, 0.015)
    cam = make_camera(synth.look_at_camera(np.array([0.2, 0.1, 1.0]), 0.7), synth.focal_from_fov(128, 30.0))
    (rgba, depth, steps), (r_rgba, r_depth, r_steps), _ = _render_pair(ctx, orc, cam, 128, 128, 0, 0)
    assert np.array_equal(steps, r_steps)
    p = _psnr(rgba[:, :3], r_rgba[:, :3])
    print(f"precision {prec}: PSNR vs FP32 oracle {p:.1f} dB")
    assert p >= 45.0


# ---------------------------------------------------------------- render buffer
@pytest.mark.parametrize("curve,srgb", [(0, 1), (1, 1), (2, 0), (3, 1)])
def test_render_frame_accumulates_like_the_oracle(small_avatar, curve, srgb):
    """Testbed::render_frame over 4 sample indices (jittered cameras): the
    shaded, accumulated and tonemapped frame against the oracle's render of
    each sample index followed by its render-buffer stages.  No early
    termination, so every ray's samples are pure march geometry."""
    torch = _torch()
    from insta_amd.capi import default_frame_config
    from oracle import oracle as O
    ctx, orc = _scene(small_config(min_transmittance=0.0), small_avatar)
    ctx.build_occupancy(0, 0.05)
    orc.build_occupancy(0, 0.05)
    res = 48
    cam = make_camera(synth.look_at_camera(np.array([0.2, -0.1, 1.0]), 1.0), synth.focal_from_fov(res, 30.0),
                      snap=False)
    fc = default_frame_config(exposure=0.3, background=(0.2, 0.5, 0.9, 1.0), tonemap_curve=curve, output_srgb=srgb)
    accum = torch.zeros((res * res, 4), dtype=torch.float32, device="cuda")
    depth = torch.empty(res * res, dtype=torch.float32, device="cuda")
    r_accum = np.zeros((res * res, 4), np.float32)
    for spp in range(4):
        out = ctx.render_frame(cam, res, res, 0, 0, spp, fc, accum, depth=depth)
        torch.cuda.synchronize()
        r_rgba, r_depth, _, _ = orc.render(cam, res, res, 0, 0, spp)
        r_out, r_dout = O.frame_buffer(r_rgba, r_depth, spp, fc, r_accum)
        assert_rel_close(out.cpu().numpy(), r_out, RTOL, f"frame spp {spp}")
        assert_rel_close(accum.cpu().numpy(), r_accum, RTOL, f"accumulation buffer spp {spp}")
        d, rd = depth.cpu().numpy(), r_dout
        assert np.array_equal(d == 0, rd == 0)
        assert_rel_close(d[rd > 0], rd[rd > 0], RTOL, "depth buffer")


def test_frame_buffer_stage_on_identical_inputs():
    """insta_frame_buffer on one composited image sequence: the same values as
    the oracle's stage to rounding of pow (GPU and host libm)."""
    torch = _torch()
    from insta_amd.capi import default_frame_config, frame_buffer
    from oracle import oracle as O
    ensure_oracle()
    rng = np.random.default_rng(12)
    n = 5000
    fc = default_frame_config(exposure=-0.5, background=(0.9, 0.1, 0.3, 0.6), tonemap_curve=2)
    accum = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    out = torch.empty_like(accum)
    r_accum = np.zeros((n, 4), np.float32)
    for spp in range(5):
        a = rng.uniform(0, 1, n).astype(np.float32)
        rgba = np.concatenate([rng.uniform(0, 1, (n, 3)).astype(np.float32) * a[:, None], a[:, None]], axis=1)
        frame_buffer(torch.from_numpy(rgba).cuda(), None, spp, fc, accum, out)
        torch.cuda.synchronize()
        r_out, _ = O.frame_buffer(rgba, None, spp, fc, r_accum)
        assert_rel_close(out.cpu().numpy(), r_out, 2e-6, f"frame buffer spp {spp}")


# ---------------------------------------------------------------- reference configuration
def test_render_with_the_reference_main_config(small_avatar, tmp_path):
    """A configs/nerf/main.json-shaped network (hash grid 16 levels x 8
    features, T = 2^18; density MLP with 2 hidden layers, colour MLP with 3;
    16 expression values) read through insta_load_config_json renders on the
    per-sample FP32 path like the oracle."""
    torch = _torch()
    import json
    from insta_amd.capi import load_config_json
    from tests.test_config_json import MAIN
    path = tmp_path / "main.json"
    path.write_text(json.dumps(MAIN))
    cfg, _ = load_config_json(path)
    for k, v in dict(grid_size=64, dt=1.0 / 256.0, max_steps=256, max_frames=2, max_triangles=4096,
                     max_vertices=4096).items():
        setattr(cfg, k, v)
    assert cfg.n_features_per_level == 8 and cfg.density_hidden_layers == 2 and cfg.n_extra_dims == 16
    ctx, orc = _scene(cfg, small_avatar)
    vals = np.random.default_rng(3).standard_normal((2, 16)).astype(np.float32)
    ctx.set_frame_extra_dims(0, torch.from_numpy(vals).cuda())
    orc.set_frame_extra_dims(0, vals)
    ctx.build_occupancy(0, 0.05)
    orc.build_occupancy(0, 0.05)
    cam = make_camera(synth.look_at_camera(np.array([0.3, 0.2, 1.0]), 1.0), synth.focal_from_fov(32, 30.0))
    (rgba, depth, steps), (r_rgba, r_depth, r_steps), _ = _render_pair(ctx, orc, cam, 32, 32, 0, 0)
    assert steps.sum() > 1000
    assert_render_matches(rgba, steps, r_rgba, r_steps, RTOL, cfg.min_transmittance, "main.json network")
