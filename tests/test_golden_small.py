"""The committed small-workload fixture (tests/golden/small_scalar.npz,
written by tests/golden/make_small_golden.py from the oracle) against the
oracle built now: the inputs regenerate to the recorded digest and every
output is reproduced -- ids, step counts, occupancy bits and table indices
exactly, floating-point outputs to 1e-6 relative (another compiler or
optimisation level may reorder nothing here, the oracle's arithmetic order
is fixed, so they agree bit for bit in practice)."""
import numpy as np
import pytest

from tests.conftest import ROOT, assert_rel_close, ensure_oracle
from tests.golden import make_small_golden as G

FIXTURE = ROOT / "tests" / "golden" / "small_scalar.npz"


@pytest.fixture(scope="module")
def fixture():
    return dict(np.load(FIXTURE, allow_pickle=False))


@pytest.fixture(scope="module")
def live():
    ensure_oracle()
    from oracle.oracle import Oracle
    cfg = G.config()
    orc = Oracle(cfg)
    av, params, pts, coords, enc_pts, cam = G.inputs(cfg, orc.layout)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    orc.set_frame(0, av.frames[0])
    orc.set_params(params)
    return orc, (av, params, pts, coords, enc_pts, cam)


def test_inputs_regenerate_to_the_recorded_digest(fixture, live):
    _, inp = live
    assert str(fixture["digest"]) == G.digest(*inp)


def test_oracle_reproduces_the_fixture(fixture, live):
    orc, (av, params, pts, coords, enc_pts, cam) = live
    ids, d2 = orc.closest_brute(0, pts)
    assert np.array_equal(ids, fixture["closest_ids"])
    assert_rel_close(np.sqrt(d2), fixture["closest_dist"], 1e-6, "distance")
    bvh_ids, _ = orc.closest(0, pts)
    assert np.array_equal(bvh_ids, fixture["closest_ids"])
    assert_rel_close(orc.deform(coords, 0), fixture["deformed"], 1e-6, "deformed")
    assert np.array_equal(orc.build_occupancy(0, G.THRESHOLD), fixture["occupancy"])
    assert np.array_equal(orc.encode_indices(enc_pts), fixture["encode_indices"])
    assert_rel_close(orc.encode(enc_pts), fixture["encode"], 1e-6, "encode")
    assert_rel_close(orc.network(coords), fixture["network"], 1e-6, "network")
    rgba, depth, steps, _ = orc.render(cam, G.RES, G.RES, 0, 0)
    assert np.array_equal(steps, fixture["steps"])
    assert_rel_close(rgba, fixture["rgba"], 1e-6, "rgba")
    assert_rel_close(depth, fixture["depth"], 1e-6, "depth")


def test_fixture_is_not_degenerate(fixture):
    # distinct triangles, a visible object, occupied cells, live features
    assert len(np.unique(fixture["closest_ids"])) > 300
    assert (fixture["rgba"][:, 3] > 0.5).sum() > 300 and fixture["steps"].sum() > 100000
    assert 1000 < np.unpackbits(fixture["occupancy"]).sum() < 64 ** 3 // 2
    assert np.abs(fixture["encode"]).max() > 1e-5
