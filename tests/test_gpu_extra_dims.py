"""Extra dims of the density input (16 per-frame expression values, the
reference's exp_params 16) on the GPU against the oracle: renders through the
MFMA (FP32, FP16, BF16) and per-sample FP32 kernels, the per-sample entry
point, training gradients, the density-grid update, and the frame values that
insta_set_frame_expression derives from its coefficients.  Bars as in the
other parity tests: step counts bit-exact, FP32 within 1e-5 relative,
FP16/BF16 renders at PSNR >= 45 dB, training gradients within 1e-3 relative
L2 of FP64."""
import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import Context, default_train_config, make_camera, small_config
from tests.conftest import assert_rel_close, assert_render_matches, ensure_oracle

pytestmark = pytest.mark.gpu
RTOL = 1e-5
X = 16
N_FRAMES = 2


def _lips_camera(av, res, frame=0):
    lips = av.masks == synth.MASK_LIPS
    c = av.frames[frame][av.faces[lips]].reshape(-1, 3).mean(axis=0)
    direction = c - synth.CENTER
    return make_camera(synth.look_at_camera(direction / np.linalg.norm(direction), 1.0),
                       synth.focal_from_fov(res, 25.0))


def _pair(cfg, av, vals, params_seed=5, table_scale=0.0):
    import torch
    ensure_oracle()
    from oracle import oracle as O
    ctx, orc = Context(cfg), O.Oracle(cfg)
    ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    ctx.set_frames(0, torch.from_numpy(np.stack(av.frames).astype(np.float32)).cuda())
    for f, v in enumerate(av.frames):
        orc.set_frame(f, v)
    params = synth.network_params(params_seed, ctx.layout, cfg)
    if table_scale:
        nm = ctx.layout.n_mlp_params
        params[nm:] = np.random.default_rng(5).uniform(-table_scale, table_scale, params.size - nm).astype(np.float32)
    ctx.set_params(torch.from_numpy(params).cuda())
    orc.set_params(params)
    ctx.set_frame_extra_dims(0, torch.from_numpy(vals).cuda())
    orc.set_frame_extra_dims(0, vals)
    for f in range(len(av.frames)):
        ctx.build_occupancy(f, 0.05)
        orc.build_occupancy(f, 0.05)
    torch.cuda.synchronize()
    return ctx, orc, O, params


def _vals(seed=2):
    return np.random.default_rng(seed).standard_normal((N_FRAMES, X)).astype(np.float32)


@pytest.fixture(scope="module")
def avatar():
    return synth.small_sphere_avatar(0, N_FRAMES)


def _psnr(a, b):
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10 * np.log10(1.0 / mse)


@pytest.mark.parametrize("prec", [0, 3])  # FP32 MFMA, per-sample FP32
def test_render_matches_oracle(avatar, prec):
    res = 48
    ctx, orc, _, _ = _pair(small_config(n_extra_dims=X, max_frames=N_FRAMES, precision=prec), avatar, _vals())
    for frame in range(N_FRAMES):
        cam = _lips_camera(avatar, res, frame)
        rgba, depth, steps = (t.cpu().numpy() for t in ctx.render(cam, res, res, frame, frame, 0))
        r_rgba, r_depth, r_steps, _ = orc.render(cam, res, res, frame, frame)
        assert_render_matches(rgba, steps, r_rgba, r_steps, RTOL, ctx.cfg.min_transmittance,
                              f"rgba frame {frame} precision {prec}")
    # the values reach the image: another frame row changes it
    cam = _lips_camera(avatar, res, 0)
    a = ctx.render(cam, res, res, 0, 0, 0)[0].cpu().numpy()
    import torch
    ctx.set_frame_extra_dims(0, torch.from_numpy(_vals(9)).cuda())
    b = ctx.render(cam, res, res, 0, 0, 0)[0].cpu().numpy()
    assert np.abs(a - b).max() > 1e-3


@pytest.mark.parametrize("prec", [1, 2])
def test_render_reduced_precision_psnr(avatar, prec):
    res = 64
    ctx, orc, _, _ = _pair(small_config(n_extra_dims=X, max_frames=N_FRAMES, precision=prec), avatar, _vals())
    cam = _lips_camera(avatar, res)
    rgba, _, steps = (t.cpu().numpy() for t in ctx.render(cam, res, res, 0, 0, 0))
    r_rgba, _, r_steps, _ = orc.render(cam, res, res, 0, 0)
    assert (steps.astype(np.uint32) == r_steps).mean() >= 0.99
    p = _psnr(rgba[:, :3], r_rgba[:, :3])
    print(f"precision {prec} with extra dims: PSNR vs FP32 oracle {p:.1f} dB")
    assert p >= 45.0


@pytest.mark.parametrize("prec", [0, 1])
def test_per_sample_entry_point_takes_the_coordinate_values(avatar, prec):
    import torch
    ctx, orc, _, _ = _pair(small_config(n_extra_dims=X, max_frames=N_FRAMES, precision=prec), avatar, _vals())
    rng = np.random.default_rng(41)
    n = 20000
    c = np.zeros((n, 11 + X), np.float32)
    c[:, 0:3] = rng.uniform(0.05, 0.95, (n, 3))
    d = rng.standard_normal((n, 3))
    c[:, 8:11] = (d / np.linalg.norm(d, axis=1, keepdims=True) + 1.0) * 0.5
    c[:, 11:] = rng.standard_normal((n, X))
    got = ctx.network_inference(torch.from_numpy(c).cuda()).cpu().numpy()
    assert_rel_close(got, orc.network(c), RTOL, "network with per-sample extra dims")


def test_training_gradients_match_fp64(avatar):
    import torch
    res, n_rays, seed = 48, 2048, 5
    cfg = small_config(n_extra_dims=X, max_frames=N_FRAMES)
    ctx, orc, O, params = _pair(cfg, avatar, _vals(), params_seed=11, table_scale=0.5)
    rng = np.random.default_rng(3)
    cams = [_lips_camera(avatar, res, f) for f in range(N_FRAMES)]
    cams[1] = make_camera(synth.random_camera(rng, 1.0), synth.focal_from_fov(res, 30.0))
    images = synth.smooth_noise_images(3, N_FRAMES, res, res)
    ctx.set_training_set(torch.from_numpy(images).cuda(), cams, list(range(N_FRAMES)))
    O.set_training_set(orc, images, cams, list(range(N_FRAMES)))
    tc = default_train_config(n_rays=n_rays, max_samples=n_rays * cfg.max_steps,
                              target_samples=n_rays * cfg.max_steps, apply_update=0, seed=seed)
    st = ctx.train_step(tc)
    g = ctx.gradients().cpu().numpy().astype(np.float64)
    g64, (rep, _), counts = O.train_grads(orc, n_rays, seed)
    assert (st.n_rays_used, st.n_samples, st.n_samples_used) == (counts[0], counts[1], counts[2])
    assert abs(st.loss - rep) <= 1e-5 * rep
    din = ctx.layout.density_in
    w0, w0_64 = g[:64 * din].reshape(64, din), g64[:64 * din].reshape(64, din)
    E = 2 * cfg.n_levels
    blocks = {"W0 encoding": (w0[:, :E], w0_64[:, :E]), "W0 extra": (w0[:, E:], w0_64[:, E:]),
              "rest": (g[64 * din:], g64[64 * din:])}
    for name, (a, b) in blocks.items():
        e = float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30))
        print(f"  {name}: rel L2 err {e:.3e}  |g| {np.linalg.norm(b):.3e}")
        assert np.linalg.norm(b) > 0 and e < 1e-3, name


def test_density_grid_update_matches_oracle(avatar):
    cfg = small_config(n_extra_dims=X, max_frames=N_FRAMES)
    ctx, orc, O, _ = _pair(cfg, avatar, _vals(), params_seed=17)
    G = cfg.grid_size
    for k, (nu, nn) in enumerate([(G ** 3, 0), (G ** 3 // 4, G ** 3 // 4)]):
        grid, mean = ctx.update_density_grid(-1, N_FRAMES, decay=0.95, n_uniform=nu, n_nonuniform=nn, seed=31,
                                             want_grid=True)
        g = grid.cpu().numpy()
        rg, rmean = O.update_density_grid(orc, -1, N_FRAMES, 0.95, nu, nn, 31)
        assert np.array_equal(g == 0, rg == 0), k
        assert_rel_close(g, rg, 1e-4, f"grid after update {k}")
        assert abs(mean - rmean) <= 1e-4 * rmean


@pytest.mark.parametrize("n_bases", [50, 10])
def test_expression_sets_the_frame_values(n_bases):
    """insta_set_frame_expression: the frame's values are its first 16
    coefficients, zero-padded below 16 (the same image as setting them)."""
    import torch
    av = synth.small_sphere_avatar(0, 1, n_bases=n_bases)
    cfg = small_config(n_extra_dims=X, max_frames=1)
    ctx = Context(cfg)
    ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
    ctx.set_params(torch.from_numpy(synth.network_params(5, ctx.layout, cfg)).cuda())
    coeffs = np.random.default_rng(7).standard_normal(n_bases).astype(np.float32)
    bases = torch.from_numpy(np.ascontiguousarray(av.bases, np.float32)).cuda()
    ctx.set_frame_expression(0, bases, torch.from_numpy(coeffs).cuda())
    ctx.build_occupancy(0, 0.05)
    res = 40
    lips = av.masks == synth.MASK_LIPS
    c = av.canonical[av.faces[lips]].reshape(-1, 3).mean(axis=0) - synth.CENTER
    cam = make_camera(synth.look_at_camera(c / np.linalg.norm(c), 1.0), synth.focal_from_fov(res, 25.0))
    r1 = ctx.render(cam, res, res, 0, 0, 0)[0].cpu().numpy()
    vals = np.zeros((1, X), np.float32)
    vals[0, :min(X, n_bases)] = coeffs[:X]
    ctx.set_frame_extra_dims(0, torch.from_numpy(np.zeros((1, X), np.float32)).cuda())
    r0 = ctx.render(cam, res, res, 0, 0, 0)[0].cpu().numpy()
    ctx.set_frame_extra_dims(0, torch.from_numpy(vals).cuda())
    r2 = ctx.render(cam, res, res, 0, 0, 0)[0].cpu().numpy()
    torch.cuda.synchronize()
    assert np.array_equal(r1, r2)
    assert np.abs(r1 - r0).max() > 1e-3


def test_full_size_render_matches_oracle():
    """The full-size shapes (9,976-triangle head, L16 T2^19, 128^3 grid) with
    the 16 extra dims, FP32 MFMA path, on a view of the lips."""
    from insta_amd.capi import full_config
    av = synth.head_avatar(seed=1, n_frames=N_FRAMES)
    res = 40
    ctx, orc, _, _ = _pair(full_config(n_extra_dims=X, max_frames=N_FRAMES, precision=0), av, _vals(3))
    for f in range(N_FRAMES):
        ctx.build_occupancy(f, 0.015)
        orc.build_occupancy(f, 0.015)
    lips = av.masks == synth.MASK_LIPS
    c = av.frames[0][av.faces[lips]].reshape(-1, 3).mean(axis=0) - synth.CENTER
    cam = make_camera(synth.look_at_camera(c / np.linalg.norm(c), 0.7), synth.focal_from_fov(res, 25.0))
    rgba, _, steps = (t.cpu().numpy() for t in ctx.render(cam, res, res, 0, 0, 0))
    r_rgba, _, r_steps, _ = orc.render(cam, res, res, 0, 0)
    assert (r_steps > 0).sum() > 100 and r_rgba[:, 3].max() > 0.05
    assert_render_matches(rgba, steps, r_rgba, r_steps, RTOL, ctx.cfg.min_transmittance,
                          "full-size rgba with extra dims", max_mismatch=0.01)


def test_frame_extra_dims_errors():
    import torch
    from insta_amd.capi import InstaError
    av = synth.small_sphere_avatar(0, 1)
    plain = Context(small_config(max_frames=2))
    with pytest.raises(InstaError):  # no extra dims in this configuration
        plain.set_frame_extra_dims(0, torch.zeros((1, X), device="cuda"))
    ctx = Context(small_config(max_frames=2, n_extra_dims=X))
    ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
    with pytest.raises(InstaError):  # frames 1..2 of 2
        ctx.set_frame_extra_dims(1, torch.zeros((2, X), device="cuda"))
    ctx.set_frame_extra_dims(1, torch.zeros((1, X), device="cuda"))
    torch.cuda.synchronize()
