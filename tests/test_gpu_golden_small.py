"""The HIP library against the committed small-workload fixture
(tests/golden/small_scalar.npz, the oracle's outputs): nearest-triangle
ids, occupancy bits, hash-grid features and march step counts bit-exact;
deformed samples, network outputs and composited pixels within 1e-5
relative (pixels whose ray stops one sample apart at the termination
threshold are bounded by that sample's weight)."""
import numpy as np
import pytest

from tests.conftest import ROOT, assert_rel_close, assert_render_matches
from tests.golden import make_small_golden as G

pytestmark = pytest.mark.gpu
FIXTURE = ROOT / "tests" / "golden" / "small_scalar.npz"
RTOL = 1e-5


def test_library_matches_the_small_fixture():
    import torch
    from insta_amd.capi import Context
    fx = dict(np.load(FIXTURE, allow_pickle=False))
    cfg = G.config()
    ctx = Context(cfg)
    av, params, pts, coords, enc_pts, cam = G.inputs(cfg, ctx.layout)
    assert str(fx["digest"]) == G.digest(av, params, pts, coords, enc_pts, cam)
    ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
    ctx.set_frames(0, torch.from_numpy(av.frames[0][None].astype(np.float32)).cuda())
    ctx.set_params(torch.from_numpy(params).cuda())
    ids, dist = ctx.closest(0, torch.from_numpy(pts).cuda())
    assert np.array_equal(ids.cpu().numpy(), fx["closest_ids"])
    assert np.array_equal(dist.cpu().numpy(), fx["closest_dist"])
    g = torch.from_numpy(coords).cuda()
    ctx.deform_samples(g, 0)
    assert_rel_close(g.cpu().numpy(), fx["deformed"], RTOL, "deformed")
    assert np.array_equal(ctx.build_occupancy(0, G.THRESHOLD).cpu().numpy(), fx["occupancy"])
    assert np.array_equal(ctx.encode(torch.from_numpy(enc_pts).cuda()).cpu().numpy(), fx["encode"])
    assert_rel_close(ctx.network_inference(torch.from_numpy(coords).cuda()).cpu().numpy(), fx["network"], RTOL,
                     "network")
    rgba, depth, steps = ctx.render(cam, G.RES, G.RES, 0, 0)
    torch.cuda.synchronize()
    assert_render_matches(rgba.cpu().numpy(), steps.cpu().numpy(), fx["rgba"], fx["steps"], RTOL,
                          cfg.min_transmittance, "rgba")
