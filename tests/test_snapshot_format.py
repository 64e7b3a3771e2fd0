"""Snapshot files on the host side (no GPU): the configuration reader of the
msgpack snapshot (insta_snapshot_config) against files written with Python's
msgpack, and malformed files rejected with an error instead of a crash."""
import ctypes as C

import msgpack
import numpy as np
import pytest

from insta_amd.capi import InstaConfig, InstaError, small_config, snapshot_config

FIELDS = [name for name, _ in InstaConfig._fields_]


def _insta_map(cfg):
    return {name: getattr(cfg, name) for name in FIELDS}


def test_config_round_trip_through_python_msgpack(tmp_path):
    cfg = small_config(max_frames=7, precision=1)
    path = tmp_path / "a.msgpack"
    # float64 floats, as nlohmann's to_msgpack writes them, and extra keys
    doc = {"encoding": {"otype": "HashGrid"}, "insta": _insta_map(cfg),
           "snapshot": {"version": 1, "params_binary": b"\x00" * 8}}
    path.write_bytes(msgpack.packb(doc, use_single_float=False))
    got = snapshot_config(path)
    for name in FIELDS:
        assert getattr(got, name) == pytest.approx(getattr(cfg, name), rel=1e-7), name


def test_config_with_single_floats_and_large_maps(tmp_path):
    cfg = small_config()
    m = _insta_map(cfg)
    m.update({f"extra_{i}": i for i in range(40)})          # map16 header
    path = tmp_path / "b.msgpack"
    path.write_bytes(msgpack.packb({"insta": m, "pad": "x" * 300, "blob": b"\x01" * 70000},
                                   use_single_float=True))
    got = snapshot_config(path)
    assert got.n_levels == cfg.n_levels and got.dt == np.float32(cfg.dt)


@pytest.mark.parametrize("payload", [
    b"",                                                     # empty file
    b"\x93\x01\x02\x03",                                     # an array, not a map
    msgpack.packb({"insta": {"n_levels": 8}}),               # incomplete configuration
    msgpack.packb({"insta": {**{n: 1 for n in FIELDS}, "n_levels": -3}}),   # negative field
    msgpack.packb({"snapshot": {}})[:-1] + b"\xc1",          # reserved type code
])
def test_malformed_snapshots_are_rejected(tmp_path, payload):
    path = tmp_path / "bad.msgpack"
    path.write_bytes(payload)
    with pytest.raises(InstaError):
        snapshot_config(path)


def test_truncated_snapshot_is_rejected(tmp_path):
    full = msgpack.packb({"insta": _insta_map(small_config()), "blob": b"\x02" * 1000})
    for cut in (1, 17, len(full) // 2, len(full) - 1):
        path = tmp_path / f"t{cut}.msgpack"
        path.write_bytes(full[:cut])
        with pytest.raises(InstaError):
            snapshot_config(path)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(InstaError):
        snapshot_config(tmp_path / "absent.msgpack")
