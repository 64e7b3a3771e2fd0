"""Shared fixtures.  CPU fixtures build scenes in the oracle; GPU fixtures
build the same scenes in the HIP library through the C ABI."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from insta_amd import synth  # noqa: E402
from insta_amd.capi import small_config, full_config  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")


def ensure_oracle():
    lib = ROOT / "oracle" / "liboracle.so"
    if not lib.exists():
        import subprocess
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


@pytest.fixture(scope="session")
def small_avatar():
    return synth.small_sphere_avatar(seed=0, n_frames=2)


@pytest.fixture(scope="session")
def head_avatar():
    return synth.head_avatar(seed=1, n_frames=4)


@pytest.fixture(scope="session")
def small_oracle(small_avatar):
    ensure_oracle()
    from oracle.oracle import Oracle
    cfg = small_config()
    o = Oracle(cfg)
    o.set_canonical_mesh(small_avatar.canonical, small_avatar.faces, small_avatar.masks)
    for f, v in enumerate(small_avatar.frames):
        o.set_frame(f, v)
    o.set_params(synth.network_params(7, o.layout, cfg))
    return o


def assert_rel_close(got, ref, rtol, what=""):
    """|got - ref| <= rtol * (|ref| + mean|ref|): relative error with the
    array's mean magnitude as the floor for entries near zero."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = np.abs(ref) + np.abs(ref).mean() + 1e-30
    err = np.abs(got - ref) / scale
    worst = np.unravel_index(np.argmax(err), err.shape)
    assert err.max() <= rtol, f"{what}: max rel err {err.max():.3e} at {worst}: got {got[worst]} ref {ref[worst]}"


def assert_render_matches(rgba, steps, r_rgba, r_steps, rtol, min_transmittance, what="", max_mismatch=0.005):
    """Rendered pixels against the oracle's.  Pixels with the oracle's sample
    count match to `rtol` (assert_rel_close).  A ray whose accumulated alpha
    lands within rounding of the termination threshold 1 - min_transmittance
    may stop one sample apart: at most `max_mismatch` of the pixels, each off
    by exactly one sample, and its RGBA off by at most that sample's weight,
    alpha * T with T at the threshold (T <= min_transmittance, alpha <= 1),
    plus the same relative tolerance."""
    rgba = np.asarray(rgba, np.float64).reshape(-1, 4)
    r_rgba = np.asarray(r_rgba, np.float64).reshape(-1, 4)
    steps = np.asarray(steps).astype(np.int64).ravel()
    r_steps = np.asarray(r_steps).astype(np.int64).ravel()
    same = steps == r_steps
    assert 1.0 - same.mean() <= max_mismatch, f"{what}: {int((~same).sum())} pixels with other sample counts"
    assert_rel_close(rgba[same], r_rgba[same], rtol, what)
    off = ~same
    if off.any():
        delta = np.abs(steps[off] - r_steps[off])
        assert (delta == 1).all(), f"{what}: sample counts differ by {delta.max()} (at most 1 allowed)"
        bound = 1.02 * min_transmittance + rtol * (np.abs(r_rgba[off]) + np.abs(r_rgba).mean())
        err = np.abs(rgba[off] - r_rgba[off])
        assert (err <= bound).all(), f"{what}: a one-sample pixel differs by {err.max():.3e} (> one sample's weight)"
