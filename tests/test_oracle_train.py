"""CPU checks of the training-step oracle: the PCG32 stream and the FP64
gradient of the training objective (central finite differences)."""
import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import small_config, make_camera
from tests.conftest import ensure_oracle

ensure_oracle()
from oracle import oracle as O  # noqa: E402


def test_pcg32_known_answers():
    # pcg32 reference demo (pcg-random.org, pcg32-demo): srandom(42, 54)
    expected = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    got = [O.pcg32_next(42, 54, k) for k in range(len(expected))]
    assert got == expected, [hex(g) for g in got]


def test_pcg32_jump_ahead_matches_stepping():
    # advance(k) (the O(log k) jump polynomial) against k single LCG steps
    seq = O.pcg32_draws(1337, 0xda3e39cb94b95bdb, 5000)
    for k in (0, 1, 2, 7, 64, 999, 4096, 4999):
        assert O.pcg32_next(1337, 0xda3e39cb94b95bdb, k) == seq[k], k
    # top bits of 4096 draws are about half ones
    assert abs(int((seq[:4096] >> 31).sum()) - 2048) < 200


def _train_oracle(n_frames=2, res=32, seed=3):
    cfg = small_config(max_frames=n_frames)
    av = synth.small_sphere_avatar(seed=seed, n_frames=n_frames)
    orc = O.Oracle(cfg)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    for f, v in enumerate(av.frames):
        orc.set_frame(f, v)
        orc.build_occupancy(f, 0.05)
    params = synth.network_params(11, orc.layout, cfg)
    # larger hash features than the U(-1e-4, 1e-4) init so the table gradients
    # are not negligible next to the MLP's
    rng = np.random.default_rng(5)
    nm = orc.layout.n_mlp_params
    params[nm:] = rng.uniform(-0.5, 0.5, size=params.size - nm).astype(np.float32)
    orc.set_params(params)
    rng = np.random.default_rng(seed)
    cams = [make_camera(synth.random_camera(rng, 1.0, front_bias=False), synth.focal_from_fov(res, 30.0))
            for _ in range(n_frames)]
    O.set_training_set(orc, synth.smooth_noise_images(seed, n_frames, res, res), cams, list(range(n_frames)))
    return orc, cfg, params


def test_train_gradient_matches_finite_differences():
    orc, cfg, params = _train_oracle()
    n_rays, seed = 48, 9
    g, (rep, obj), counts = O.train_grads(orc, n_rays, seed)
    assert counts[0] > 10 and counts[1] > counts[0]
    assert obj > 0 and abs(rep - obj / 3) < 1e-12
    lay = orc.layout
    nm = lay.n_mlp_params
    rng = np.random.default_rng(0)
    # parameters with a clearly non-zero gradient in every weight block and in the tables
    bounds = np.cumsum([0] + [o * i for (o, i) in lay.mlp_shapes(cfg)] + [lay.n_grid_params])
    picks = []
    for b0, b1 in zip(bounds[:-1], bounds[1:]):
        blk = np.arange(b0, b1)
        big = blk[np.abs(g[blk]) > 0.05 * np.abs(g[blk]).max()]
        picks += list(rng.choice(big, size=min(5, len(big)), replace=False))
    worst = 0.0
    for p in picks:
        # central differences at three step sizes: a ReLU / Huber kink close to
        # the parameter value spoils the larger steps, never all three; FP64
        # noise stays far below the tolerance even at the smallest
        errs = []
        for h in (2.0 ** -17, 2.0 ** -19, 2.0 ** -21):
            pp, pm = params.copy(), params.copy()
            pp[p] = params[p] + np.float32(h)
            pm[p] = params[p] - np.float32(h)
            orc.set_params(pp)
            _, (_, lp), _ = O.train_grads(orc, n_rays, seed, want_grads=False)
            orc.set_params(pm)
            _, (_, lm), _ = O.train_grads(orc, n_rays, seed, want_grads=False)
            fd = (lp - lm) / (float(pp[p]) - float(pm[p]))
            errs.append(abs(fd - g[p]) / (abs(g[p]) + 1e-12))
        worst = max(worst, min(errs))
        assert min(errs) < 1e-5, (int(p), errs, g[p])
    orc.set_params(params)
    assert worst < 1e-5


def test_train_gradient_zero_for_untouched_table_entries():
    orc, cfg, params = _train_oracle()
    g, _, _ = O.train_grads(orc, 32, 4)
    nm = orc.layout.n_mlp_params
    tab = g[nm:]
    # only the corners of sampled cells receive gradient: most entries of the
    # fine hashed levels stay exactly zero, the dense coarse level is touched
    assert (tab == 0).mean() > 0.5
    assert np.abs(tab).max() > 0
