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


def _compact3(v):
    v = v & 0x09249249
    v = (v | (v >> 2)) & 0x030C30C3
    v = (v | (v >> 4)) & 0x0300F00F
    v = (v | (v >> 8)) & 0x030000FF
    v = (v | (v >> 16)) & 0x000003FF
    return v


def _untrained_cameras():
    """Two cameras with a non-square image and unequal focal lengths, both
    looking at the box from the front half-space, so cells behind them and
    cells off their frusta exist."""
    from insta_amd.capi import make_camera
    W, H = 48, 32
    cams = []
    for d, fov in (((0.3, 0.2, 1.0), 25.0), ((-0.5, 0.1, 0.8), 18.0)):
        c = make_camera(synth.look_at_camera(np.array(d), 0.9), synth.focal_from_fov(H, fov))
        c.focal_px[0] = c.focal_px[1] * 1.3
        cams.append(c)
    return cams, W, H


def _projection_classes(cams, W, H, G):
    """Per cell (Morton order): inside = the centre projects into some
    camera's image (pinhole, principal point at the image centre); outside =
    for every camera the centre is behind it or projects more than the
    cell's half-diagonal (in pixels at its depth) plus a pixel beyond the
    image edge."""
    i = np.arange(G ** 3, dtype=np.int64)
    c = (np.stack([_compact3(i), _compact3(i >> 1), _compact3(i >> 2)], 1) + 0.5) / G
    r = 0.5 * np.sqrt(3.0) / G
    inside = np.zeros(len(i), bool)
    outside = np.ones(len(i), bool)
    for cam in cams:
        M = np.array(cam.matrix, np.float64).reshape(3, 4)
        loc = c - M[:, 3]
        x, y, z = loc @ M[:, 0], loc @ M[:, 1], loc @ M[:, 2]
        fx, fy = cam.focal_px[0], cam.focal_px[1]
        with np.errstate(divide="ignore", invalid="ignore"):
            u, v = fx * x / z + 0.5 * W, fy * y / z + 0.5 * H
            mu, mv = fx * r / z + 1.0, fy * r / z + 1.0
        front = z > 1e-3
        inside |= front & (u > 0) & (u < W) & (v > 0) & (v < H)
        far = (u < -mu) | (u > W + mu) | (v < -mv) | (v > H + mv)
        outside &= ~front | far
    return inside, outside


def test_untrained_cells_are_those_no_camera_sees():
    """mark_untrained_density_grid (testbed_nerf.cu:391-431) against an
    independent pinhole projection: cells whose centre lands in a training
    image stay trainable, cells clearly outside every frustum are -1; -1
    cells are never occupied and stay -1; a camera added later re-marks."""
    orc, cfg, av = _scene(n_frames=2, zero_mlp=True)
    G = cfg.grid_size
    cams, W, H = _untrained_cameras()
    imgs = np.full((2, H, W, 4), 0.5, np.float32)
    O.set_training_set(orc, imgs[:1], cams[:1], [0])
    grid, _ = O.update_density_grid(orc, -1, 2, 0.95, G ** 3, 0, 5)
    inside, outside = _projection_classes(cams[:1], W, H, G)
    assert inside.sum() > 1000 and outside.sum() > 1000
    assert not (grid[inside] < 0).any()
    assert (grid[outside] == -1.0).all()
    bits = np.unpackbits(orc.get_occupancy(-1), bitorder="little").astype(bool)
    assert not bits[grid < 0].any() and bits.any()
    untrained = grid < 0
    grid2, _ = O.update_density_grid(orc, -1, 2, 0.95, G ** 3 // 4, G ** 3 // 4, 5)
    assert np.array_equal(grid2 < 0, untrained)
    # a second image (the count changes): cells it sees become trainable again
    O.set_training_set(orc, imgs, cams, [0, 1])
    grid3, _ = O.update_density_grid(orc, -1, 2, 0.95, G ** 3 // 4, G ** 3 // 4, 5)
    inside2, outside2 = _projection_classes(cams, W, H, G)
    assert not (grid3[inside2] < 0).any() and (grid3[outside2] == -1.0).all()
    assert (untrained & ~(grid3 < 0)).sum() > 100
