"""warp_random_samples and the density-grid EMA on the GPU against the
oracle: same random frames, same cells, same EMA and bitfield."""
import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import Context, small_config
from tests.conftest import assert_rel_close, ensure_oracle

pytestmark = pytest.mark.gpu


def _pair(n_frames=3):
    import torch
    ensure_oracle()
    from oracle import oracle as O
    cfg = small_config(max_frames=n_frames)
    av = synth.small_sphere_avatar(seed=6, n_frames=n_frames)
    params = synth.network_params(17, O.Oracle(cfg).layout, cfg)
    ctx = Context(cfg)
    ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
    ctx.set_frames(0, torch.from_numpy(np.stack(av.frames)).cuda())
    ctx.set_params(torch.from_numpy(params).cuda())
    orc = O.Oracle(cfg)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    for f, v in enumerate(av.frames):
        orc.set_frame(f, v)
    orc.set_params(params)
    torch.cuda.synchronize()
    return ctx, orc, O, cfg


def test_random_frame_positions_match_oracle():
    import torch
    ctx, orc, O, cfg = _pair()
    rng = np.random.default_rng(1)
    pts = rng.uniform(0.1, 0.9, (20000, 3)).astype(np.float32)
    dev = torch.from_numpy(pts).cuda()
    ctx.deform_positions(dev, 3, seed=99, rng_offset=12345)
    got = dev.cpu().numpy()
    ref = O.deform_positions(orc, pts, 3, 99, 12345)
    assert_rel_close(got, ref, 1e-5, "canonical positions")


def test_density_grid_updates_match_oracle():
    import torch
    ctx, orc, O, cfg = _pair()
    G = cfg.grid_size
    schedule = [(G ** 3, 0), (G ** 3 // 4, G ** 3 // 4), (G ** 3 // 4, G ** 3 // 4)]
    for k, (nu, nn) in enumerate(schedule):
        grid, mean = ctx.update_density_grid(-1, 3, decay=0.95, n_uniform=nu, n_nonuniform=nn, seed=31,
                                             want_grid=True)
        g = grid.cpu().numpy()
        rg, rmean = O.update_density_grid(orc, -1, 3, 0.95, nu, nn, 31)
        # same cells sampled: the zero pattern matches exactly
        assert np.array_equal(g == 0, rg == 0), k
        assert_rel_close(g, rg, 1e-4, f"grid after update {k}")
        assert abs(mean - rmean) <= 1e-4 * rmean
    # the march slot: bits agree except cells within rounding of the threshold
    got_bits = np.unpackbits(ctx.get_occupancy(-1).cpu().numpy(), bitorder="little").astype(bool)
    ref_bits = np.unpackbits(orc.get_occupancy(-1), bitorder="little").astype(bool)
    assert np.array_equal(got_bits, g > min(0.01, mean))
    assert (got_bits == ref_bits).mean() > 0.999
    # a render through the EMA slot (canonical mesh, slot -1)
    from insta_amd.capi import make_camera
    cam = make_camera(synth.look_at_camera(np.array([0.2, 0.3, 1.0]), 1.0), synth.focal_from_fov(32, 30.0))
    rgba, _, steps = ctx.render(cam, 32, 32, -1, -1, 0)
    torch.cuda.synchronize()
    assert int(steps.sum().item()) > 0


def test_untrained_cells_match_oracle():
    """With a training set, the first update marks the cells no training
    camera sees (-1, mark_untrained_density_grid) exactly as the oracle;
    they stay out of the samples, the EMA and the bitfield, and a training
    set with another image count re-marks."""
    import torch
    from tests.test_oracle_density import _untrained_cameras
    ctx, orc, O, cfg = _pair()
    G = cfg.grid_size
    cams, W, H = _untrained_cameras()
    imgs = np.full((2, H, W, 4), 0.5, np.float32)
    for n in (1, 2):
        ctx.set_training_set(torch.from_numpy(imgs[:n]).cuda(), cams[:n], list(range(n)))
        O.set_training_set(orc, imgs[:n], cams[:n], list(range(n)))
        for k, (nu, nn) in enumerate([(G ** 3, 0), (G ** 3 // 4, G ** 3 // 4)]):
            grid, mean = ctx.update_density_grid(-1, 3, decay=0.95, n_uniform=nu, n_nonuniform=nn, seed=31,
                                                 want_grid=True)
            g = grid.cpu().numpy()
            rg, rmean = O.update_density_grid(orc, -1, 3, 0.95, nu, nn, 31)
            assert np.array_equal(g < 0, rg < 0), (n, k)
            assert (rg < 0).sum() > 1000 and (rg >= 0).sum() > 1000
            assert np.array_equal(g == 0, rg == 0), (n, k)
            assert_rel_close(g, rg, 1e-4, f"grid {n} {k}")
            assert abs(mean - rmean) <= 1e-4 * rmean
            bits = np.unpackbits(ctx.get_occupancy(-1).cpu().numpy(), bitorder="little").astype(bool)
            assert not bits[g < 0].any()
