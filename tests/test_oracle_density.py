"""CPU checks of the oracle's warp_random_samples and density-grid update."""
import numpy as np

from insta_amd import synth
from insta_amd.capi import small_config
from tests.conftest import ensure_oracle

ensure_oracle()
from oracle import oracle as O  # noqa: E402


def _scene(n_frames=3, zero_mlp=False):
    cfg = small_config(max_frames=n_frames)
    av = synth.small_sphere_avatar(seed=4, n_frames=n_frames)
    orc = O.Oracle(cfg)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    for f, v in enumerate(av.frames):
        orc.set_frame(f, v)
    params = synth.network_params(13, orc.layout, cfg)
    if zero_mlp:
        params[: orc.layout.n_mlp_params] = 0.0
    orc.set_params(params)
    return orc, cfg, av


def test_random_frame_warp_uses_frames_but_the_last():
    orc, cfg, av = _scene(n_frames=3)
    rng = np.random.default_rng(0)
    pts = (0.5 + rng.uniform(-0.3, 0.3, (4000, 3))).astype(np.float32)
    got = O.deform_positions(orc, pts, n_frames=3, seed=21)
    # per-frame reference: the AoS warp of each point through each frame
    coords = np.zeros((len(pts), 11), np.float32)
    coords[:, :3] = pts
    coords[:, 4] = 0
    coords[:, 8:11] = 0.5
    per_frame = [orc.deform(coords, frame=f)[:, :3] for f in range(3)]
    match = np.stack([np.all(got == pf, axis=1) for pf in per_frame], axis=1)
    assert match[:, :2].any(axis=1).all()          # every point went through frame 0 or 1
    assert not (match[:, 2] & ~match[:, :2].any(axis=1)).any()  # frame 2 never alone
    used = match[:, :2].sum(axis=0)
    assert used.min() > 1000                        # both frames drawn


def test_density_grid_ema_closed_form():
    # zero MLP weights: density logit 0, optical thickness exp(bias) * dt everywhere
    orc, cfg, _ = _scene(n_frames=2, zero_mlp=True)
    G = cfg.grid_size
    # update 1: G^3 uniform samples hit every cell once (odd multiplier mod 2^k)
    grid, mean = O.update_density_grid(orc, -1, 2, 0.5, G ** 3, 0, 5)
    t = grid[0]
    assert abs(float(t) - np.exp(cfg.density_logit_bias) * cfg.dt) < 1e-6 * float(t)
    assert np.all(grid == t)
    assert abs(mean - t) < 1e-6 * t
    # update 2: a few nonuniform samples; untouched cells decay by half
    grid2, mean2 = O.update_density_grid(orc, -1, 2, 0.5, 0, 1000, 5)
    hit = grid2 == t
    assert np.all(grid2[~hit] == np.float32(t * np.float32(0.5)))
    assert 500 < hit.sum() <= 1000
    expect_mean = (hit.sum() * float(t) + (~hit).sum() * float(t) * 0.5) / G ** 3
    assert abs(mean2 - expect_mean) < 1e-5 * expect_mean
    # bitfield: only cells above min(0.01, mean) = mean, i.e. the re-hit cells
    assert mean2 < 0.01
    bits = np.unpackbits(orc.get_occupancy(-1), bitorder="little").astype(bool)
    assert np.array_equal(bits, hit)
