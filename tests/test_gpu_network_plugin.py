"""The network plugin entry points (NerfNetwork forward_impl / backward_impl /
density, nerf_network.h:238-414) through the C ABI against the oracle:
forward outputs to 1e-5 of the FP32 oracle network (tables of the workload's
U(-1e-4, 1e-4)), and with trained-scale tables (U(-0.5, 0.5)) the parameter
gradients per weight block and the hash tables, and the position/direction
input gradients, to 1e-3 relative L2 of the FP64 oracle; density outputs to
1e-5 of the FP32 oracle and of FP64 (per column, relative L2)."""
import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import InstaError, small_config
from tests.conftest import assert_rel_close, ensure_oracle

pytestmark = pytest.mark.gpu


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _pair(n_extra, table_scale=None, seed=7):
    torch = _torch()
    ensure_oracle()
    from insta_amd.capi import Context
    from oracle.oracle import Oracle
    cfg = small_config(n_extra_dims=n_extra)
    ctx, orc = Context(cfg), Oracle(cfg)
    p = synth.network_params(seed, ctx.layout, cfg)
    if table_scale:
        p[ctx.layout.n_mlp_params:] = np.random.default_rng(seed).uniform(-table_scale, table_scale,
                                                                           ctx.layout.n_grid_params)
    ctx.set_params(torch.from_numpy(p).cuda())
    orc.set_params(p)
    return ctx, orc, cfg


def _coords(rng, n, n_extra):
    c = np.zeros((n, 11 + n_extra), np.float32)
    c[:, 0:3] = rng.uniform(0.15, 0.85, (n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c[:, 8:11] = (d + 1) * 0.5
    if n_extra:
        c[:, 11:] = rng.standard_normal((n, n_extra))
    return c


def _rel(a, b):
    return float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30))


@pytest.mark.parametrize("n_extra", [0, 16])
def test_forward_matches_oracle(n_extra):
    torch = _torch()
    ctx, orc, _ = _pair(n_extra)
    c = _coords(np.random.default_rng(1), 20000, n_extra)
    out = ctx.network_forward(torch.from_numpy(c).cuda())
    torch.cuda.synchronize()
    assert_rel_close(out.cpu().numpy(), orc.network(c), 1e-5, "plugin forward")


@pytest.mark.parametrize("n_extra", [0, 16])
def test_backward_matches_fp64(n_extra):
    torch = _torch()
    from oracle import oracle as O
    ctx, orc, cfg = _pair(n_extra, table_scale=0.5)
    rng = np.random.default_rng(2)
    n = 3000
    c = _coords(rng, n, n_extra)
    dL = rng.standard_normal((n, 4)).astype(np.float32) * 1e-2
    cd = torch.from_numpy(c).cuda()
    out = ctx.network_forward(cd)
    dinput = torch.zeros_like(cd)
    g = ctx.network_backward(cd, torch.from_numpy(dL).cuda(), dL_dinput_dev=dinput)
    torch.cuda.synchronize()
    ref_out, ref_g, ref_di = O.network_plugin_d(orc, c.astype(np.float64), dL.astype(np.float64))
    assert_rel_close(out.cpu().numpy(), ref_out, 2e-4, "plugin forward vs FP64 (trained-scale tables)")
    g = g.cpu().numpy().astype(np.float64)
    lay = ctx.layout
    b = np.cumsum([0] + [o * i for (o, i) in lay.mlp_shapes(cfg)] + [lay.n_grid_params])
    for k, (b0, b1) in enumerate(zip(b[:-1], b[1:])):
        assert _rel(g[b0:b1], ref_g[b0:b1]) < 1e-3, (k, _rel(g[b0:b1], ref_g[b0:b1]))
    di = dinput.cpu().numpy().astype(np.float64)
    assert _rel(di[:, 0:3], ref_di[:, 0:3]) < 1e-3
    assert _rel(di[:, 8:11], ref_di[:, 3:6]) < 1e-3
    # rows the reference does not write are left alone
    assert (di[:, 3:8] == 0).all() and (di[:, 11:] == 0).all()


def test_backward_accumulates_and_checks_its_forward():
    torch = _torch()
    ctx, _, _ = _pair(0, table_scale=0.5)
    rng = np.random.default_rng(3)
    c = torch.from_numpy(_coords(rng, 1000, 0)).cuda()
    dL = torch.from_numpy(rng.standard_normal((1000, 4)).astype(np.float32)).cuda()
    with pytest.raises(InstaError):   # no stored forward yet
        ctx.network_backward(c, dL)
    ctx.network_forward(c)
    g1 = ctx.network_backward(c, dL)
    g2 = ctx.network_backward(c, dL, grads_dev=g1.clone(), accumulate=True)
    torch.cuda.synchronize()
    # the hash-table gradients are FP32 atomic sums whose order varies from
    # launch to launch, so the second pass agrees to rounding, not bit for bit
    a, b = g2.cpu().numpy(), 2 * g1.cpu().numpy()
    assert_rel_close(a, b, 1e-5, "accumulated gradients")
    assert _rel(a, b) < 1e-6
    with pytest.raises(InstaError):   # another batch size than the stored forward
        ctx.network_backward(c[:500], dL[:500])


@pytest.mark.parametrize("n_extra", [0, 16])
def test_density_matches_fp64(n_extra):
    torch = _torch()
    from oracle import oracle as O
    ctx, orc, _ = _pair(n_extra, table_scale=0.05)
    pos = np.random.default_rng(4).uniform(0.0, 1.0, (20000, 3)).astype(np.float32)
    out = ctx.network_density(torch.from_numpy(pos).cuda())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    # the FP32 restatement to 1e-5 relative
    assert_rel_close(got, O.network_density(orc, pos), 1e-5, "density outputs vs the FP32 oracle")
    # FP64: outputs that cancel to near zero carry FP32 rounding of the
    # 64-term sums, so each column to 1e-5 in relative L2, entries to 1e-4
    ref = O.network_density_d(orc, pos.astype(np.float64))
    for k in range(16):
        assert _rel(got[:, k], ref[:, k]) < 1e-5, (k, _rel(got[:, k], ref[:, k]))
    assert_rel_close(got, ref, 1e-4, "density outputs vs FP64")
