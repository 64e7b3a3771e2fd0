"""Snapshot save/load on the GPU: a trained avatar written with
insta_save_snapshot and read back into a fresh context renders the same image,
carries the same parameters, density grid (FP16, as the reference stores it)
and optimizer state, and continues training like the original; the file is
plain msgpack with the reference's snapshot keys."""
import msgpack
import numpy as np
import pytest

from insta_amd.capi import Context, InstaError, default_train_config, snapshot_config, small_config
from tests.test_gpu_train import N_FRAMES, _gpu, _scene

pytestmark = pytest.mark.gpu


def _trained(steps=3):
    import torch
    scene = _scene()
    ctx, _ = _gpu(scene)
    # a batch cap above the composited count: compaction keeps every sample, so
    # the batch does not depend on the order waves reserve their slots
    tc = default_train_config(n_rays=1024, max_samples=1 << 18, target_samples=1 << 18, seed=9)
    for _ in range(steps):
        ctx.train_step(tc)
    # density grid EMA into the canonical occupancy slot (the training
    # marches read the frames' own slots)
    grid, _ = ctx.update_density_grid(-1, N_FRAMES, n_uniform=64 ** 3, want_grid=True)
    torch.cuda.synchronize()
    return scene, ctx, tc, grid.cpu().numpy()


def test_snapshot_round_trip(tmp_path):
    import torch
    scene, ctx, tc, grid = _trained()
    path = tmp_path / "avatar.msgpack"
    ctx.save_snapshot(path, include_optimizer=True)

    # the file: msgpack with the reference's keys
    doc = msgpack.unpackb(path.read_bytes(), raw=False)
    snap = doc["snapshot"]
    for key in ("version", "params_type", "params_binary", "density_grid_size", "density_grid_binary",
                "training_step", "aabb", "render_aabb", "nerf"):
        assert key in snap, key
    assert doc["encoding"]["otype"] == "HashGrid" and doc["encoding"]["n_levels"] == ctx.cfg.n_levels
    params = ctx.params().cpu().numpy()
    assert snap["params_type"] == "float" and snap["n_params"] == params.size
    assert np.frombuffer(snap["params_binary"], np.float32).tobytes() == params.tobytes()
    g16 = np.frombuffer(snap["density_grid_binary"], np.float16)
    assert g16.size == 64 ** 3 and np.array_equal(g16, grid.astype(np.float16))
    assert snap["training_step"] == 3 and snap["optimizer"]["step"] == 3
    # the training cameras as NerfDataset (json_binding.h): focal length,
    # principal point, resolution and a 3x4 start/end transform per image
    cfg_, av_, cams_, images_, _ = scene
    ds = snap["nerf"]["dataset"]
    assert ds["n_images"] == N_FRAMES and len(ds["metadata"]) == N_FRAMES and len(ds["xforms"]) == N_FRAMES
    for i, cam in enumerate(cams_):
        md = ds["metadata"][i]
        assert md["resolution"] == [images_.shape[2], images_.shape[1]]
        assert np.allclose(md["focal_length"], list(cam.focal_px)) and md["camera_distortion"] is None
        assert np.allclose(md["principal_point"], list(cam.screen_center))
        m = np.asarray(ds["xforms"][i]["start"], np.float32)
        assert m.shape == (3, 4) and np.array_equal(m.ravel(), np.asarray(list(cam.matrix), np.float32))
        assert ds["xforms"][i]["end"] == ds["xforms"][i]["start"]
    assert set(ds["render_aabb"]) == {"min", "max"} and ds["aabb_scale"] == 1
    cfg2 = snapshot_config(path)
    assert cfg2.n_levels == ctx.cfg.n_levels and cfg2.grid_size == ctx.cfg.grid_size

    # a fresh context (same mesh and training set, other parameters) loads it
    other = np.random.default_rng(1).uniform(-1e-3, 1e-3, params.size).astype(np.float32)
    ctx2, _ = _gpu(scene, params=other)
    ctx2.load_snapshot(path, occupancy_slot=-1)
    torch.cuda.synchronize()
    assert np.array_equal(ctx2.params().cpu().numpy(), params)
    # the density grid's bitfield was rebuilt from the FP16 grid: threshold
    # min(0.01, mean) (the device mean is a fixed-order FP32 sum: cells within
    # rounding of the threshold may differ)
    g32 = g16.astype(np.float32)
    thr = np.float32(min(0.01, float(g32.astype(np.float64).mean())))
    bits = np.packbits((g32 > thr).reshape(-1, 8), axis=1, bitorder="little").ravel()
    got = ctx2.get_occupancy(-1).cpu().numpy()
    assert (np.unpackbits(got) == np.unpackbits(bits)).mean() > 0.9999

    # same image from both contexts (frame 0, its own mesh occupancy)
    from insta_amd import synth
    from insta_amd.capi import make_camera
    cam = make_camera(synth.random_camera(np.random.default_rng(4), 1.0, front_bias=False),
                      synth.focal_from_fov(64, 30.0))
    a = [t.clone() for t in ctx.render(cam, 64, 64, 0, 0)]
    b = ctx2.render(cam, 64, 64, 0, 0)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)

    # training continues from the same optimizer state and ray counters
    s1 = ctx.train_step(tc)
    s2 = ctx2.train_step(tc)
    torch.cuda.synchronize()
    # Testbed::load_snapshot resumes at training_step + 1 (testbed.cu:3011);
    # Adam continues from the saved moments and step
    assert s1.step == 4 and s2.step == 5 and s1.adam_step == s2.adam_step == 4
    assert s1.n_rays == s2.n_rays and s1.n_samples == s2.n_samples
    assert s1.n_samples_used == s2.n_samples_used and s1.loss == pytest.approx(s2.loss, rel=1e-6)
    g1, g2 = ctx.gradients().cpu().numpy(), ctx2.gradients().cpu().numpy()
    nm = ctx.layout.n_mlp_params
    # the batch order (wave reservations) and the tables' FP32 atomics vary
    # between runs: gradients agree to rounding (Adam with eps 1e-15 would turn
    # sign noise of near-zero gradients into full steps, so compare gradients)
    for sl in (slice(0, nm), slice(nm, None)):
        assert np.linalg.norm(g1[sl] - g2[sl]) <= 1e-5 * np.linalg.norm(g1[sl])


def test_reload_without_optimizer_state_keeps_the_training_step(tmp_path):
    """Without saved moments Adam restarts (bias correction from step 1)
    while the training step -- which times the density-grid updates of the
    training loop -- resumes at the saved value + 1, as the reference's
    load_snapshot does whatever the optimizer state (testbed.cu:3011)."""
    import torch
    scene, ctx, tc, _ = _trained(steps=4)
    path = tmp_path / "no_moments.msgpack"
    ctx.save_snapshot(path)
    doc = msgpack.unpackb(path.read_bytes(), raw=False)
    assert doc["snapshot"]["training_step"] == 4 and "optimizer" not in doc["snapshot"]
    ctx2, _ = _gpu(scene)
    ctx2.load_snapshot(path, occupancy_slot=-1)
    s = ctx2.train_step(tc)
    torch.cuda.synchronize()
    assert s.step == 6 and s.adam_step == 1
    # the loop's schedule reads the training step: at step 6 an update is due
    # every clamp(6 / 16, 1, 16) = 1 steps, past the 256-step warm-up it
    # would be every 16; the library's own loop keeps counting from there
    last = ctx2.train_loop(tc, 2, density_slot=-1, n_frames=N_FRAMES)
    torch.cuda.synchronize()
    assert last.step == 8 and last.adam_step == 3


def test_load_rejects_another_network_shape(tmp_path):
    scene, ctx, _, _ = _trained(steps=1)
    path = tmp_path / "a.msgpack"
    ctx.save_snapshot(path)
    doc = msgpack.unpackb(path.read_bytes(), raw=False)
    assert "optimizer" not in doc["snapshot"]
    wrong = Context(small_config(max_frames=N_FRAMES, log2_hashmap_size=12))
    with pytest.raises(InstaError):
        wrong.load_snapshot(path)
    with pytest.raises(InstaError):
        ctx.load_snapshot(tmp_path / "missing.msgpack")
