"""C-ABI surface checks that need no GPU: the library loads, exports every
function include/insta_capi.h declares, and its host-only entry points
(configuration defaults, layout, validation) behave."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

from insta_amd import capi
from tests.conftest import ensure_oracle

ROOT = Path(__file__).resolve().parents[1]


def _declared_functions():
    text = (ROOT / "include" / "insta_capi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(insta_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    bound = {n for n, _, _ in capi.SIGNATURES}
    assert bound == set(declared)


def test_default_config_is_the_full_size_workload():
    cfg = capi.default_config()
    assert (cfg.n_levels, cfg.n_features_per_level, cfg.log2_hashmap_size, cfg.base_resolution) == (16, 2, 19, 16)
    assert math.isclose(cfg.per_level_scale, 2 ** (7 / 15), rel_tol=1e-6)
    assert cfg.grid_size == 128 and cfg.max_steps == 1024
    assert math.isclose(cfg.dt, math.sqrt(3) / 1024, rel_tol=1e-6)


@pytest.mark.parametrize("make", [capi.small_config, capi.full_config])
def test_layout_matches_oracle(make):
    ensure_oracle()
    from oracle.oracle import Oracle
    cfg = make()
    lay = capi.query_layout(cfg)
    orc = Oracle(cfg).layout
    for f, _ in capi.InstaLayout._fields_:
        a, b = getattr(lay, f), getattr(orc, f)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), f
        else:
            assert a == b, f
    # finest level of the full-size grid is 2048 cells across
    if cfg.n_levels == 16:
        assert lay.level_resolution[15] in (2048, 2049)
        assert lay.n_grid_params > 16 * (1 << 19)


def test_invalid_configs_are_rejected():
    lib = capi.load_library()
    lay = capi.InstaLayout()
    for bad in [dict(n_levels=0), dict(n_features_per_level=3), dict(grid_size=100), dict(max_triangles=20000),
                dict(sh_degree=3), dict(min_transmittance=1.0), dict(dt=0.0), dict(n_extra_dims=8),
                dict(n_extra_dims=17)]:
        cfg = capi.full_config(**bad)
        st = lib.insta_query_layout(C.byref(cfg), C.byref(lay))
        assert st == 1, bad
        assert lib.insta_last_error()


def test_extra_dims_layout():
    # 16 expression values appended to the density input: W0 grows by 64 x 16
    plain, extra = capi.query_layout(capi.full_config()), capi.query_layout(capi.full_config(n_extra_dims=16))
    assert extra.density_in == plain.density_in + 16 == 48
    assert extra.n_mlp_params == plain.n_mlp_params + 64 * 16
    assert extra.n_grid_params == plain.n_grid_params


def test_null_context_is_an_error_not_a_crash():
    lib = capi.load_library()
    assert lib.insta_set_frame_extra_dims(None, 0, 0, None, None) == 1
    assert lib.insta_deform_samples(None, None, 0, 11, -1, None) == 1
    assert lib.insta_render(None, None, 1, 1, 0, 0, 0, None, None, None, None) == 1
    assert b"null" in lib.insta_last_error()


def test_synthetic_mesh_counts():
    from insta_amd import synth
    v, f = synth.octasphere(4)
    assert (len(v), len(f)) == (1026, 2048)
    u, f2 = synth.head_ellipsoid()
    assert (len(u), len(f2)) == (5023, 9976)
    assert len(np.unique(f2)) == 5023
