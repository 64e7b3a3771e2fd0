"""Workspace reservation: insta_render, insta_train_step and the network
plugin never allocate; they run inside what insta_reserve set aside and
fail with INSTA_ERR_STATE (naming insta_reserve) beyond it.  Calls through
the C ABI directly, without the Python wrapper's own reservations."""
import ctypes as C

import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import (InstaError, InstaTrainStats, _ptr, check, default_train_config, make_camera,
                            small_config)

pytestmark = pytest.mark.gpu
STATE = 4


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _ctx():
    torch = _torch()
    from insta_amd.capi import Context
    cfg = small_config(max_frames=4)
    av = synth.small_sphere_avatar(seed=0, n_frames=4)
    ctx = Context(cfg)
    ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
    ctx.set_frames(0, torch.from_numpy(np.stack(av.frames)).cuda())
    ctx.set_params(torch.from_numpy(synth.network_params(7, ctx.layout, cfg)).cuda())
    for f in range(4):
        ctx.build_occupancy(f, 0.05)
    return ctx, av


def _raw_render(ctx, w, h):
    torch = _torch()
    cam = make_camera(synth.look_at_camera(np.array([0.3, 0.2, 1.0]), 1.0), synth.focal_from_fov(w, 30.0))
    rgba = torch.empty((w * h, 4), dtype=torch.float32, device="cuda")
    check(ctx.lib.insta_render(ctx.h, C.byref(cam), w, h, 0, 0, 0, _ptr(rgba), None, None, ctx.stream()),
          "insta_render")
    return rgba


def test_render_runs_inside_its_reservation():
    torch = _torch()
    ctx, _ = _ctx()
    with pytest.raises(InstaError) as e:       # nothing reserved yet
        _raw_render(ctx, 32, 32)
    assert e.value.status == STATE and "insta_reserve" in str(e.value)
    ctx.reserve(render_pixels=48 * 48)
    _raw_render(ctx, 48, 48)                   # warm: module loads, first launches
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    a = _raw_render(ctx, 48, 48)
    b = _raw_render(ctx, 40, 32)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    assert torch.isfinite(a).all() and torch.isfinite(b).all()
    with pytest.raises(InstaError) as e:       # one pixel beyond
        _raw_render(ctx, 48 * 48 + 1, 1)
    assert e.value.status == STATE


def test_train_step_runs_inside_its_reservation():
    torch = _torch()
    ctx, av = _ctx()
    rng = np.random.default_rng(1)
    cams = [make_camera(synth.random_camera(rng, 1.0, front_bias=False), synth.focal_from_fov(64, 30.0))
            for _ in range(4)]
    ctx.set_training_set(torch.from_numpy(synth.smooth_noise_images(5, 4, 64, 64)).cuda(), cams, [0, 1, 2, 3])
    tc = default_train_config(n_rays=1024, max_samples=1 << 18, target_samples=1 << 15, seed=5)

    def raw_step():
        st = InstaTrainStats()
        check(ctx.lib.insta_train_step(ctx.h, C.byref(tc), C.byref(st), ctx.stream()), "insta_train_step")
        return st

    with pytest.raises(InstaError) as e:
        raw_step()
    assert e.value.status == STATE and "insta_reserve" in str(e.value)
    ctx.reserve(train_rays=1024, train_samples=1 << 18, train_batch=1 << 15)
    raw_step()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    st = raw_step()
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0 and st.n_samples > 0
    tc.n_rays = 2048                           # more rays than reserved
    with pytest.raises(InstaError) as e:
        raw_step()
    assert e.value.status == STATE
