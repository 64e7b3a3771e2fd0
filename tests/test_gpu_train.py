"""Training step on the GPU against the FP64 oracle (same rays, same
samples): sample counts bit-exact, loss and gradients within tolerance,
FP16 grid-gradient accumulation with loss scaling, Adam's closed form, a
decreasing loss, and renders after training against the oracle."""
import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import Context, default_train_config, make_camera, small_config
from tests.conftest import assert_rel_close, assert_render_matches, ensure_oracle

pytestmark = pytest.mark.gpu

N_FRAMES, RES, SEED = 4, 64, 3


def _scene(table_scale=0.5, **cfg_kw):
    cfg = small_config(max_frames=N_FRAMES, **cfg_kw)
    av = synth.small_sphere_avatar(seed=SEED, n_frames=N_FRAMES)
    rng = np.random.default_rng(SEED)
    cams = [make_camera(synth.random_camera(rng, 1.0, front_bias=False), synth.focal_from_fov(RES, 30.0))
            for _ in range(N_FRAMES)]
    images = synth.smooth_noise_images(SEED, N_FRAMES, RES, RES)
    return cfg, av, cams, images, table_scale


def _params(layout, cfg, table_scale):
    p = synth.network_params(11, layout, cfg)
    if table_scale:
        nm = layout.n_mlp_params
        p[nm:] = np.random.default_rng(5).uniform(-table_scale, table_scale, p.size - nm).astype(np.float32)
    return p


def _gpu(scene, params=None):
    import torch
    cfg, av, cams, images, ts = scene
    ctx = Context(cfg)
    ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
    ctx.set_frames(0, torch.from_numpy(np.stack(av.frames)).cuda())
    if params is None:
        params = _params(ctx.layout, cfg, ts)
    ctx.set_params(torch.from_numpy(params).cuda())
    for f in range(N_FRAMES):
        ctx.build_occupancy(f, 0.05)
    ctx.set_training_set(torch.from_numpy(images).cuda(), cams, list(range(N_FRAMES)))
    torch.cuda.synchronize()
    return ctx, params


def _oracle(scene, params):
    ensure_oracle()
    from oracle import oracle as O
    cfg, av, cams, images, _ = scene
    orc = O.Oracle(cfg)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    for f, v in enumerate(av.frames):
        orc.set_frame(f, v)
        orc.build_occupancy(f, 0.05)
    orc.set_params(params)
    O.set_training_set(orc, images, cams, list(range(N_FRAMES)))
    return orc, O


def _blocks(layout, cfg):
    b = np.cumsum([0] + [o * i for (o, i) in layout.mlp_shapes(cfg)] + [layout.n_grid_params])
    names = ["W0", "W1", "W2", "W3", "W4", "tables"]
    return list(zip(names, b[:-1], b[1:]))


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30))


@pytest.fixture(scope="module")
def scene():
    return _scene()


@pytest.mark.parametrize("n_rays,seed", [(1024, 5), (4096, 6)])
def test_batch_and_gradients_match_fp64(scene, n_rays, seed):
    import torch
    ctx, params = _gpu(scene)
    cfg = scene[0]
    # batch capacity for every composited sample: nothing truncated, as in the oracle
    tc = default_train_config(n_rays=n_rays, max_samples=n_rays * cfg.max_steps,
                              target_samples=n_rays * cfg.max_steps, apply_update=0, seed=seed)
    st = ctx.train_step(tc)
    g = ctx.gradients().cpu().numpy().astype(np.float64)
    torch.cuda.synchronize()
    orc, O = _oracle(scene, params)
    g64, (rep, obj), counts = O.train_grads(orc, n_rays, seed)
    print(f"rays {st.n_rays_used}/{counts[0]} samples {st.n_samples}/{counts[1]} used {st.n_samples_used}/{counts[2]} "
          f"loss {st.loss:.6e}/{rep:.6e}")
    errs = {}
    for name, b0, b1 in _blocks(ctx.layout, cfg):
        errs[name] = _rel_l2(g[b0:b1], g64[b0:b1])
        print(f"  {name}: rel L2 err {errs[name]:.3e}  |g| {np.linalg.norm(g64[b0:b1]):.3e}")
    assert (st.n_rays_used, st.n_samples) == (counts[0], counts[1])   # march bit-exact
    assert st.n_samples_used == counts[2]
    # FP32 encode, MLPs, compositing and backward against FP64 on identical canonical samples
    assert abs(st.loss - rep) <= 1e-5 * rep
    # FP32 sums over ~1e5 samples with cancellation, and ReLU masks that flip
    # between FP32 and FP64 for near-zero pre-activations (observed <= 3e-4)
    for name, e in errs.items():
        assert e < 1e-3, name


def test_fp16_grid_gradients_with_loss_scale(scene):
    ctx, params = _gpu(scene)
    cfg = scene[0]
    n_rays, seed = 2048, 7
    tc = default_train_config(n_rays=n_rays, max_samples=n_rays * cfg.max_steps,
                              target_samples=n_rays * cfg.max_steps, apply_update=0, seed=seed,
                              grid_grad_fp16=1, loss_scale=128.0)
    ctx.train_step(tc)
    g = ctx.gradients().cpu().numpy().astype(np.float64)
    orc, O = _oracle(scene, params)
    g64, _, _ = O.train_grads(orc, n_rays, seed)
    for name, b0, b1 in _blocks(ctx.layout, cfg):
        e = _rel_l2(g[b0:b1], g64[b0:b1])
        print(f"  {name}: rel L2 err {e:.3e}")
        # MLP gradients stay FP32 (loss scale divides out exactly: a power of
        # two); hash-grid gradients carry FP16 accumulation error: each add
        # rounds to 11 bits of the running sum of its chunk of the batch
        # (8192 samples per partial table; observed 1.4e-3)
        assert e < (1e-2 if name == "tables" else 1e-3), name


@pytest.fixture(scope="module")
def mixed_scene():
    # T = 2^16: levels 0-1 (<= 16384 entries) take the LDS partial-table
    # scatter, levels 2-7 the per-sample scatter into the per-XCD copies
    return _scene(log2_hashmap_size=16)


@pytest.mark.parametrize("fp16", [0, 1])
def test_mixed_scatter_levels_match_fp64(mixed_scene, fp16):
    ctx, params = _gpu(mixed_scene)
    cfg = mixed_scene[0]
    sizes = [ctx.layout.level_size[l] for l in range(cfg.n_levels)]
    assert sizes[1] <= 16384 < sizes[2], sizes
    n_rays, seed = 2048, 9
    kw = dict(grid_grad_fp16=1, loss_scale=128.0) if fp16 else {}
    tc = default_train_config(n_rays=n_rays, max_samples=n_rays * cfg.max_steps,
                              target_samples=n_rays * cfg.max_steps, apply_update=0, seed=seed, **kw)
    ctx.train_step(tc)
    g = ctx.gradients().cpu().numpy().astype(np.float64)
    orc, O = _oracle(mixed_scene, params)
    g64, _, _ = O.train_grads(orc, n_rays, seed)
    nm = ctx.layout.n_mlp_params
    off = [ctx.layout.level_offset[l] for l in range(cfg.n_levels)] + [ctx.layout.n_grid_params // 2]
    rng = lambda l0, l1: slice(nm + 2 * off[l0], nm + 2 * off[l1])
    errs = [_rel_l2(g[rng(l, l + 1)], g64[rng(l, l + 1)]) for l in range(cfg.n_levels)]
    for l, e in enumerate(errs):
        print(f"  level {l} ({sizes[l]} entries): rel L2 err {e:.3e}")
    lds = _rel_l2(g[rng(0, 2)], g64[rng(0, 2)])
    per_sample = _rel_l2(g[rng(2, cfg.n_levels)], g64[rng(2, cfg.n_levels)])
    print(f"  LDS levels {lds:.3e}, per-sample levels {per_sample:.3e}")
    # FP32 against FP64 on identical samples: summation order, and ReLU masks
    # that flip for near-zero pre-activations, which on a fine level move the
    # few entries such a sample touches (observed <= 1.4e-3 on one level);
    # FP16: 11-bit running sums (per chunk on the LDS levels, per XCD copy on
    # the others).  A dropped chunk, copy or level offset errs by O(1).
    assert max(lds, per_sample) < (2e-2 if fp16 else 1e-3)
    assert max(errs) < (5e-2 if fp16 else 5e-3)
    for name, b0, b1 in _blocks(ctx.layout, cfg)[:-1]:
        assert _rel_l2(g[b0:b1], g64[b0:b1]) < 1e-3, name


def test_adam_first_step_closed_form(scene):
    import torch
    ctx, params = _gpu(scene)
    cfg = scene[0]
    tc = default_train_config(n_rays=1024, max_samples=1024 * cfg.max_steps, target_samples=1024 * cfg.max_steps,
                              seed=8)
    ctx.train_step(tc)
    g = ctx.gradients().cpu().numpy()
    p1 = ctx.params().cpu().numpy()
    torch.cuda.synchronize()
    f32 = np.float32
    b1, b2, lr, eps = f32(tc.beta1), f32(tc.beta2), f32(tc.learning_rate), f32(tc.epsilon)
    m = b1 * f32(0) + (f32(1) - b1) * g
    v = b2 * f32(0) + (f32(1) - b2) * (g * g)
    bc1 = f32(1.0 - float(tc.beta1))
    bc2 = f32(1.0 - float(tc.beta2))
    expect = params - lr * (m / bc1) / (np.sqrt(v / bc2) + eps)
    np.testing.assert_allclose(p1, expect, rtol=2e-6, atol=1e-9)


def test_compaction_caps_the_batch(scene):
    # a batch smaller than the composited samples: the step still runs, the
    # sample counts are unchanged and the adaptive ray count shrinks
    ctx, _ = _gpu(scene)
    cfg = scene[0]
    tc = default_train_config(n_rays=4096, max_samples=4096 * cfg.max_steps, target_samples=1 << 14, seed=6,
                              apply_update=0)
    st = ctx.train_step(tc)
    assert st.n_samples_used > (1 << 14)
    g = ctx.gradients().cpu().numpy()
    assert np.isfinite(g).all() and np.abs(g).max() > 0


def test_training_reduces_loss(scene):
    ctx, _ = _gpu(_scene(table_scale=0.0))
    tc = default_train_config(target_samples=1 << 15, max_samples=1 << 18, seed=9)
    losses = [ctx.train_step(tc).loss for _ in range(60)]
    print("loss", losses[0], "->", losses[-1])
    assert np.mean(losses[-5:]) < 0.5 * np.mean(losses[:3])


def test_render_after_training_matches_oracle(scene):
    import torch
    ctx, _ = _gpu(scene)
    cfg = scene[0]
    tc = default_train_config(target_samples=1 << 14, max_samples=1 << 18, seed=10)
    for _ in range(5):
        ctx.train_step(tc)
    trained = ctx.params().cpu().numpy()
    orc, O = _oracle(scene, trained)
    cam = scene[2][1]
    rgba, depth, steps = ctx.render(cam, 48, 48, 1, 1, 0)
    torch.cuda.synchronize()
    o_rgba, o_depth, o_steps, _ = orc.render(cam, 48, 48, 1, 1, 0)
    assert_render_matches(rgba.cpu().numpy(), steps.cpu().numpy(), o_rgba, o_steps, 1e-4, ctx.cfg.min_transmittance,
                          "rgba after training")


def test_one_occupancy_slot_for_every_ray_matches_fp64(scene):
    """Rays marching one named grid (here the canonical mesh's, as the
    reference's single density grid) instead of their frame's own: counts
    bit-exact and gradients as with per-frame grids."""
    import torch
    ctx, params = _gpu(scene)
    cfg = scene[0]
    ctx.build_occupancy(-1, 0.05)
    n_rays, seed = 2048, 12
    tc = default_train_config(n_rays=n_rays, max_samples=n_rays * cfg.max_steps,
                              target_samples=n_rays * cfg.max_steps, apply_update=0, seed=seed, occupancy_slot=-1)
    st = ctx.train_step(tc)
    g = ctx.gradients().cpu().numpy().astype(np.float64)
    torch.cuda.synchronize()
    orc, O = _oracle(scene, params)
    orc.build_occupancy(-1, 0.05)
    O.set_training_occupancy(orc, -1)
    g64, (rep, _), counts = O.train_grads(orc, n_rays, seed)
    assert (st.n_rays_used, st.n_samples, st.n_samples_used) == counts
    assert abs(st.loss - rep) <= 1e-5 * rep
    for name, b0, b1 in _blocks(ctx.layout, cfg):
        assert _rel_l2(g[b0:b1], g64[b0:b1]) < 1e-3, name
    # the per-frame grids give another batch (the check above is not vacuous)
    O.set_training_occupancy(orc, -2)
    assert O.train_grads(orc, n_rays, seed, want_grads=False)[2] != counts


def test_train_loop_follows_the_reference_schedule(scene):
    """insta_train_loop = density-grid updates before steps s with
    s % clamp(s / 16, 1, 16) == 0 (G^3 uniform samples below step 256) and
    training steps marching that grid: the same as the calls made by hand."""
    import torch
    n_steps = 40
    cfg = scene[0]
    G3 = cfg.grid_size ** 3
    tc = default_train_config(n_rays=1024, max_samples=1 << 18, target_samples=1 << 18, seed=21)
    a, _ = _gpu(scene)
    st_a = a.train_loop(tc, n_steps, density_slot=-1, n_frames=N_FRAMES, decay=0.95, seed=5)
    b, _ = _gpu(scene)
    tb = default_train_config(n_rays=1024, max_samples=1 << 18, target_samples=1 << 18, seed=21, occupancy_slot=-1)
    updates = 0
    for s in range(n_steps):
        if s % min(max(s // 16, 1), 16) == 0:
            b.update_density_grid(-1, N_FRAMES, decay=0.95, n_uniform=G3, seed=5)
            updates += 1
        st_b = b.train_step(tb)
    torch.cuda.synchronize()
    assert updates == 36 and st_a.step == st_b.step == n_steps
    # FP32 atomics add in another order: the runs agree to rounding
    assert st_a.loss == pytest.approx(st_b.loss, rel=1e-3)
    occ_a, occ_b = a.get_occupancy(-1).cpu().numpy(), b.get_occupancy(-1).cpu().numpy()
    assert (np.unpackbits(occ_a) == np.unpackbits(occ_b)).mean() > 0.999
    assert np.unpackbits(occ_a).sum() > 0


def test_train_loop_density_schedule_matches_oracle(scene):
    """insta_train_loop's density-grid updates against the oracle's update
    called on the reference schedule by hand (before step s when
    s % clamp(s / 16, 1, 16) == 0, G^3 uniform samples below step 256).
    Learning rate 0 keeps the parameters fixed, so the grid depends only on
    which updates ran, in which order and with which sample counts: the
    march slot's bits agree with the oracle's."""
    import torch
    n_steps = 40
    cfg = scene[0]
    G3 = cfg.grid_size ** 3
    ctx, params = _gpu(scene)
    tc = default_train_config(n_rays=1024, max_samples=1 << 18, target_samples=1 << 18, seed=21, learning_rate=0.0)
    st = ctx.train_loop(tc, n_steps, density_slot=-1, n_frames=N_FRAMES, decay=0.95, seed=5)
    torch.cuda.synchronize()
    assert st.step == n_steps
    assert np.array_equal(ctx.params().cpu().numpy(), params)
    orc, O = _oracle(scene, params)
    updates = 0
    for s in range(n_steps):
        if s % min(max(s // 16, 1), 16) == 0:
            O.update_density_grid(orc, -1, N_FRAMES, 0.95, G3, 0, 5)
            updates += 1
    assert updates == 36
    got = np.unpackbits(ctx.get_occupancy(-1).cpu().numpy(), bitorder="little").astype(bool)
    ref = np.unpackbits(orc.get_occupancy(-1), bitorder="little").astype(bool)
    assert ref.sum() > 1000 and (~ref).sum() > 1000
    assert (got == ref).mean() > 0.999


def test_adam_l2_term_against_the_fp64_restatement(scene):
    """Two steps with l2_reg: the parameters follow the oracle's FP64 Adam
    with the L2 term (moments carried over), given the step's own loss
    gradients (which exclude the L2 term)."""
    import torch
    from oracle.oracle import adam_update_d
    ctx, params = _gpu(scene)
    cfg = scene[0]
    tc = default_train_config(n_rays=1024, max_samples=1024 * cfg.max_steps, target_samples=1024 * cfg.max_steps,
                              seed=8, l2_reg=0.05)
    w = params.astype(np.float64)
    m = v = np.zeros_like(w)
    gs = []
    for t in (1, 2):
        ctx.train_step(tc)
        g = ctx.gradients().cpu().numpy().astype(np.float64)
        gs.append(g)
        got = ctx.params().cpu().numpy().astype(np.float64)
        torch.cuda.synchronize()
        w, m, v = adam_update_d(w, g, m, v, t, tc.learning_rate, tc.beta1, tc.beta2, tc.epsilon, tc.l2_reg)
        # FP32 arithmetic against FP64 on the same gradients
        np.testing.assert_allclose(got, w, rtol=1e-5, atol=1e-7)
        w = got  # continue from the device's weights
    # the L2 term moved weights whose loss gradient was zero (untouched table entries)
    idle = (gs[0] == 0) & (gs[1] == 0) & (np.abs(params) > 0.05)
    assert idle.any() and (np.abs(got[idle]) < np.abs(params[idle])).all()


def test_ema_average_is_what_inference_uses(scene):
    """Three steps under the Ema optimizer: the inference parameters are the
    FP64 debiased average of the training weights after each step, the
    training weights stay Adam's, and a render uses the average (it equals
    the render of a context given the average as its parameters)."""
    import torch
    from insta_amd.capi import Context
    from oracle.oracle import ema_update_d
    ctx, params = _gpu(scene)
    cfg = scene[0]
    decay = 0.9
    tc = default_train_config(n_rays=1024, max_samples=1024 * cfg.max_steps, target_samples=1024 * cfg.max_steps,
                              seed=9, ema_decay=decay)
    a = np.zeros(ctx.layout.n_params)
    for t in (1, 2, 3):
        ctx.train_step(tc)
        w = ctx.params().cpu().numpy().astype(np.float64)
        inf = ctx.inference_params().cpu().numpy().astype(np.float64)
        torch.cuda.synchronize()
        a, deb = ema_update_d(a, w, decay, t)
        np.testing.assert_allclose(inf, deb, rtol=1e-5, atol=1e-7)
        assert not np.array_equal(inf, w)
    # a render through the average, against a context holding it as weights
    av = scene[1]
    cam = make_camera(synth.look_at_camera(np.array([0.2, 0.1, 1.0]), 1.0), synth.focal_from_fov(48, 30.0))
    ctx.build_occupancy(0, 0.05)
    r1 = [x.clone() for x in ctx.render(cam, 48, 48, 0, 0)]
    ref = Context(cfg)
    ref.set_canonical_mesh(av.canonical, av.faces, av.masks)
    ref.set_frames(0, torch.from_numpy(np.stack(av.frames)).cuda())
    ref.set_params(ctx.inference_params())
    ref.build_occupancy(0, 0.05)
    r2 = ref.render(cam, 48, 48, 0, 0)
    torch.cuda.synchronize()
    for x, y in zip(r1, r2):
        assert torch.equal(x, y)
    # new weights end the averaging: inference reads them directly
    ctx.set_params(torch.from_numpy(params).cuda())
    assert torch.equal(ctx.inference_params().cpu(), torch.from_numpy(params))


def _supervision(scene, seed=21):
    """Mask ids covering every rule (weighted 120/130/140, unsupervised <= 90
    and > 250, plain) and depth images around the sphere's distance, a fifth
    of the texels without depth."""
    cfg, av, cams, images, _ = scene
    rng = np.random.default_rng(seed)
    shape = (N_FRAMES, RES, RES)
    seg = rng.choice(np.array([0, 90, 91, 120, 130, 140, 200, 250, 251], np.uint8), size=shape)
    depth = rng.uniform(0.4, 1.3, shape).astype(np.float32)
    depth[rng.uniform(size=shape) < 0.2] = 0.0
    return seg, depth


@pytest.mark.parametrize("with_seg", [True, False])
def test_seg_and_depth_supervision_match_fp64(scene, with_seg):
    """Segmentation and depth terms of the loss: gradients against the FP64
    oracle on the same rays, the reported loss unchanged by them, the
    gradients clearly different from the unsupervised ones, and a new
    training set without supervision."""
    import torch
    ctx, params = _gpu(scene)
    cfg = scene[0]
    seg, depth = _supervision(scene)
    n_rays, seed = 2048, 12
    tc = default_train_config(n_rays=n_rays, max_samples=n_rays * cfg.max_steps,
                              target_samples=n_rays * cfg.max_steps, apply_update=0, seed=seed,
                              seg_weight=4.0, depth_weight=2.0)
    ctx.set_training_supervision(torch.from_numpy(seg).cuda() if with_seg else None, torch.from_numpy(depth).cuda())
    st = ctx.train_step(tc)
    g = ctx.gradients().cpu().numpy().astype(np.float64)
    torch.cuda.synchronize()
    orc, O = _oracle(scene, params)
    g_plain, (rep_plain, _), _ = O.train_grads(orc, n_rays, seed)
    O.set_training_supervision(orc, seg if with_seg else None, depth, 4.0, 2.0)
    g64, (rep, obj), counts = O.train_grads(orc, n_rays, seed)
    assert (st.n_rays_used, st.n_samples, st.n_samples_used) == counts
    # the colour loss is reported without the weights (testbed_nerf.cu:1500-1507;
    # the oracle's sum over rays runs in threads: equal to rounding)
    assert rep == pytest.approx(rep_plain, rel=1e-12)
    assert abs(st.loss - rep) <= 1e-5 * rep
    for name, b0, b1 in _blocks(ctx.layout, cfg):
        e = _rel_l2(g[b0:b1], g64[b0:b1])
        moved = _rel_l2(g_plain[b0:b1], g64[b0:b1])
        print(f"  {name}: rel L2 err {e:.3e}, unsupervised gradient {moved:.3e}")
        assert e < 1e-3, name
        if with_seg or name in ("W0", "W1", "tables"):
            assert moved > 20 * e, name
        else:
            # depth alone reaches the density logit only: the colour MLP's
            # gradients are the unsupervised ones
            assert moved < 1e-12, name
    # a new training set starts without supervision
    _, av, cams, images, _ = scene
    ctx.set_training_set(torch.from_numpy(images).cuda(), cams, list(range(N_FRAMES)))
    ctx.train_step(tc)
    g2 = ctx.gradients().cpu().numpy().astype(np.float64)
    for name, b0, b1 in _blocks(ctx.layout, cfg):
        assert _rel_l2(g2[b0:b1], g_plain[b0:b1]) < 1e-3, name
