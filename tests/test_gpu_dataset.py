"""A capture read by insta_amd.dataset drives the device path: its cameras
render the same images as the cameras it was written from, and it trains."""
import json

import numpy as np
import pytest

from insta_amd import dataset, synth
from insta_amd.capi import Context, default_train_config, make_camera, small_config
from tests.test_dataset import _write_obj

pytestmark = pytest.mark.gpu


def test_capture_renders_and_trains(tmp_path):
    import torch
    from PIL import Image
    av = synth.small_sphere_avatar(seed=6, n_frames=2)
    _write_obj(tmp_path / "canonical.obj", av.canonical - 0.5, av.faces)
    rng = np.random.default_rng(1)
    res, focal = 48, synth.focal_from_fov(48, 30.0)
    cams, frames = [], []
    for i, v in enumerate(av.frames):
        _write_obj(tmp_path / f"m{i}.obj", v - 0.5, av.faces)
        Image.fromarray(rng.integers(0, 256, (res, res, 4), dtype=np.uint8), "RGBA").save(tmp_path / f"i{i}.png")
        m = synth.random_camera(rng, 1.0, front_bias=False).reshape(3, 4).astype(np.float64)
        cams.append(make_camera(m, focal, snap=False))
        t = np.eye(4)
        t[:3, :4] = m
        t[:3, 3] -= 0.5                     # the loader adds the scene offset
        frames.append({"file_path": f"i{i}.png", "mesh_path": f"m{i}.obj", "transform_matrix": t.tolist()})
    (tmp_path / "transforms.json").write_text(json.dumps({"fl_x": focal, "w": res, "h": res, "frames": frames}))
    ds = dataset.load_transforms(tmp_path / "transforms.json")
    ctx = Context(small_config(max_frames=2))
    ctx.set_canonical_mesh(ds.canonical, ds.faces, ds.masks)
    ctx.set_frames(0, torch.from_numpy(np.stack(ds.frames)).cuda())
    ctx.set_params(torch.from_numpy(synth.network_params(3, ctx.layout, ctx.cfg)).cuda())
    for f in range(2):
        ctx.build_occupancy(f, 0.05)
    a = [t.clone() for t in ctx.render(ds.cameras[1], res, res, 1, 1)]
    b = ctx.render(cams[1], res, res, 1, 1)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.allclose(x, y, rtol=0, atol=1e-6)
    assert int(a[2].sum()) > 0                 # the sphere is in view
    ctx.set_training_set(torch.from_numpy(ds.images()).cuda(), ds.cameras, [0, 1])
    st = ctx.train_step(default_train_config(n_rays=512, seed=2))
    torch.cuda.synchronize()
    assert st.n_samples > 0 and np.isfinite(st.loss)


def test_trainer_reloads_subsets_with_supervision(tmp_path):
    """insta_amd.training.Trainer over a four-image capture with masks and
    depth: subsets of two images, reloaded every 5 steps (3 loads in 12
    steps), the steps counted by the context, the loss finite throughout."""
    import torch
    from PIL import Image
    from insta_amd.training import Trainer
    av = synth.small_sphere_avatar(seed=6, n_frames=4)
    _write_obj(tmp_path / "canonical.obj", av.canonical - 0.5, av.faces)
    rng = np.random.default_rng(2)
    res, focal = 32, synth.focal_from_fov(32, 30.0)
    frames = []
    for i, v in enumerate(av.frames):
        _write_obj(tmp_path / f"m{i}.obj", v - 0.5, av.faces)
        Image.fromarray(rng.integers(0, 256, (res, res, 4), dtype=np.uint8), "RGBA").save(tmp_path / f"i{i}.png")
        seg = np.zeros((res, res, 4), np.uint8)
        seg[..., 0] = rng.choice([0, 100, 120, 200], (res, res))
        Image.fromarray(seg, "RGBA").save(tmp_path / f"s{i}.png")
        Image.fromarray(rng.integers(500, 1500, (res, res)).astype(np.uint16)).save(tmp_path / f"d{i}.png")
        m = synth.random_camera(rng, 1.0, front_bias=False).reshape(3, 4).astype(np.float64)
        t = np.eye(4)
        t[:3, :4] = m
        t[:3, 3] -= 0.5
        frames.append({"file_path": f"i{i}.png", "mesh_path": f"m{i}.obj", "seg_mask_path": f"s{i}.png",
                       "depth_path": f"d{i}.png", "transform_matrix": t.tolist()})
    (tmp_path / "transforms.json").write_text(json.dumps(
        {"fl_x": focal, "w": res, "h": res, "integer_depth_scale": 0.001, "frames": frames}))
    ds = dataset.load_transforms(tmp_path / "transforms.json")
    assert ds.seg_masks() is not None and ds.depth_images() is not None
    ctx = Context(small_config(max_frames=2))
    ctx.set_canonical_mesh(ds.canonical, ds.faces, ds.masks)
    ctx.set_params(torch.from_numpy(synth.network_params(3, ctx.layout, ctx.cfg)).cuda())
    tc = default_train_config(n_rays=512, max_samples=1 << 16, target_samples=1 << 16, seed=2,
                              seg_weight=40.0, depth_weight=1.25)
    tr = Trainer(ctx, ds, tc, images_to_load=2, reload_every=5, frame_occupancy_threshold=0.05)
    subsets = []
    for n in (3, 4, 5):                        # chunks that straddle the reloads at 5 and 10
        st = tr.train(n)
        torch.cuda.synchronize()
        subsets.append(list(tr.subset))
        assert np.isfinite(st.loss) and st.n_samples > 0
    assert tr.step_count == st.step == 12
    assert tr.loads == 3
    for sub in subsets:
        assert len(sub) == 2 and set(sub) <= {0, 1, 2, 3}
