"""The network plugin's FP64 restatement (forward, parameter and input
gradients, density) in the oracle, pinned by the oracle's FP32 network
(another code path), by central finite differences of the FP64 forward for
every input gradient and a sample of parameter gradients, and by the density
logit of the forward with extra dims 1.0."""
import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import small_config
from tests.conftest import assert_rel_close, ensure_oracle


def _oracle(n_extra=0, table_scale=0.5, seed=7):
    ensure_oracle()
    from oracle.oracle import Oracle
    cfg = small_config(n_extra_dims=n_extra)
    o = Oracle(cfg)
    p = synth.network_params(seed, o.layout, cfg)
    # tables well away from zero so the encoding carries the gradients
    rng = np.random.default_rng(seed)
    p[o.layout.n_mlp_params:] = rng.uniform(-table_scale, table_scale, o.layout.n_grid_params)
    o.set_params(p)
    return o, cfg


def _coords(rng, n, n_extra):
    c = np.zeros((n, 11 + n_extra), np.float64)
    c[:, 0:3] = rng.uniform(0.2, 0.8, (n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c[:, 8:11] = (d + 1) * 0.5
    if n_extra:
        c[:, 11:] = rng.standard_normal((n, n_extra))
    return c


@pytest.mark.parametrize("n_extra", [0, 16])
def test_forward_matches_the_fp32_network(n_extra):
    from oracle import oracle as O
    o, cfg = _oracle(n_extra)
    c = _coords(np.random.default_rng(1), 300, n_extra)
    out, _, _ = O.network_plugin_d(o, c)
    ref = o.network(c.astype(np.float32)).astype(np.float64)
    # FP32 rounding through four layers with tables of U(-0.5, 0.5)
    assert_rel_close(out, ref, 2e-4, "FP64 forward vs FP32 network")


@pytest.mark.parametrize("n_extra", [0, 16])
def test_input_gradients_match_finite_differences(n_extra):
    from oracle import oracle as O
    o, cfg = _oracle(n_extra)
    rng = np.random.default_rng(2)
    c = _coords(rng, 24, n_extra)
    dL = rng.standard_normal((24, 4))
    _, _, di = O.network_plugin_d(o, c, dL)
    h = 1e-7
    fd = np.zeros_like(di)
    for k, col in enumerate((0, 1, 2, 8, 9, 10)):
        cp, cm = c.copy(), c.copy()
        cp[:, col] += h
        cm[:, col] -= h
        op, _, _ = O.network_plugin_d(o, cp)
        om, _, _ = O.network_plugin_d(o, cm)
        fd[:, k] = ((op - om) * dL).sum(axis=1) / (2 * h)
    # a step can cross a cell face or a ReLU kink: compare per sample, most
    # samples agree to FD accuracy
    err = np.abs(fd - di) / (np.abs(di) + np.abs(di).mean())
    good = (err < 1e-5).all(axis=1)
    assert good.mean() > 0.9, err.max(axis=1)
    assert np.abs(di[:, :3]).max() > 1e-3 and np.abs(di[:, 3:]).max() > 1e-3


def test_parameter_gradients_match_finite_differences():
    from oracle import oracle as O
    o, cfg = _oracle(16)
    rng = np.random.default_rng(3)
    c = _coords(rng, 16, 16)
    dL = rng.standard_normal((16, 4))
    _, g, _ = O.network_plugin_d(o, c, dL)
    lay = o.layout
    params = synth.network_params(7, lay, cfg)
    params[lay.n_mlp_params:] = np.random.default_rng(7).uniform(-0.5, 0.5, lay.n_grid_params)
    bounds = np.cumsum([0] + [a * b for (a, b) in lay.mlp_shapes(cfg)])
    picks = [int(rng.integers(bounds[k], bounds[k + 1])) for k in range(len(bounds) - 1)]
    picks += [int(p) for p in np.flatnonzero(np.abs(g[lay.n_mlp_params:]) > 1e-3)[:6] + lay.n_mlp_params]
    for idx in picks:
        best = np.inf
        for h in (1e-2, 1e-3, 1e-4):   # float32 parameters: the largest step that stays linear wins
            pp, pm = params.copy(), params.copy()
            pp[idx] += h
            pm[idx] -= h
            o.set_params(pp)
            op, _, _ = O.network_plugin_d(o, c)
            o.set_params(pm)
            om, _, _ = O.network_plugin_d(o, c)
            fd = ((op - om) * dL).sum() / (float(pp[idx]) - float(pm[idx]))
            best = min(best, abs(fd - g[idx]) / (abs(g[idx]) + 1e-6))
        o.set_params(params)
        assert best < 1e-3, (idx, best)


def test_density_is_the_forward_density_with_extras_one():
    from oracle import oracle as O
    o, cfg = _oracle(16)
    rng = np.random.default_rng(4)
    c = _coords(rng, 200, 16)
    c[:, 11:] = 1.0
    out, _, _ = O.network_plugin_d(o, c)
    dens = O.network_density_d(o, c[:, 0:3])
    assert dens.shape == (200, 16)
    assert np.abs(dens[:, 0] - out[:, 3]).max() < 1e-12


@pytest.mark.parametrize("n_extra", [0, 16])
def test_fp32_density_is_the_fp32_network_density(n_extra):
    """The FP32 density restatement: column 0 is the FP32 network's density
    logit bit for bit (same sums, same order; extra dims 1.0), and every
    column is the FP64 density to FP32 rounding."""
    from oracle import oracle as O
    o, cfg = _oracle(n_extra, table_scale=0.05)
    c = _coords(np.random.default_rng(5), 2000, n_extra).astype(np.float32)
    if n_extra:
        c[:, 11:] = 1.0
    dens = O.network_density(o, c[:, 0:3])
    assert dens.dtype == np.float32 and dens.shape == (2000, 16)
    assert np.array_equal(dens[:, 0], o.network(c)[:, 3])
    ref = O.network_density_d(o, c[:, 0:3].astype(np.float64))
    for k in range(16):
        assert np.linalg.norm(dens[:, k] - ref[:, k]) <= 1e-5 * np.linalg.norm(ref[:, k])
