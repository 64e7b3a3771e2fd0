"""The request loop (insta_server_*): a served request equals a direct render
of the same expression and camera, and the oracle's render; results are the
same alone and in a concurrent mix; admission respects the HBM budget."""
import numpy as np
import pytest

from insta_amd import synth
from insta_amd.capi import (PRECISION_FP16, PRECISION_FP32, Context, InstaError, Server, default_server_config,
                            make_camera, small_config)
from tests.conftest import assert_rel_close, assert_render_matches, ensure_oracle

pytestmark = pytest.mark.gpu


def _avatar_ctx(seed, n_frames, precision=PRECISION_FP32, n_extra=0):
    import torch
    av = synth.small_sphere_avatar(seed=seed, n_frames=0, n_bases=50)
    cfg = small_config(max_frames=n_frames, precision=precision, n_extra_dims=n_extra)
    ctx = Context(cfg)
    ctx.set_canonical_mesh(av.canonical, av.faces, av.masks)
    params = synth.network_params(100 + seed, ctx.layout, cfg)
    ctx.set_params(torch.from_numpy(params).cuda())
    bases = torch.from_numpy(av.bases).cuda()
    torch.cuda.synchronize()
    return av, ctx, bases, params


def _request(rng, res, fov=30.0):
    cam = make_camera(synth.random_camera(rng, 1.0, front_bias=False), synth.focal_from_fov(res, fov))
    return rng.normal(0.0, 1.0, 50).astype(np.float32), cam


def _serve(server, reqs):
    """reqs: list of (avatar, coeffs, cam, res); returns images (host) in order."""
    import torch
    outs, tickets = [], []
    for aid, coeffs, cam, res in reqs:
        out = torch.zeros((res * res, 4), dtype=torch.float32, device="cuda")
        tickets.append(server.submit(aid, coeffs, cam, res, res, out, tag=len(outs)))
        outs.append(out)
    results = {}
    while len(results) < len(reqs):
        for r in server.poll(wait=True):
            results[r.user_tag] = r
    assert all(r.status == 0 for r in results.values())
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs], results


def test_served_request_matches_direct_render_and_oracle():
    import torch
    av, ctx, bases, params = _avatar_ctx(1, n_frames=3)
    rng = np.random.default_rng(5)
    coeffs, cam = _request(rng, 48)
    server = Server(default_server_config(n_slots=2, hbm_budget_bytes=1 << 30))
    aid = server.add_avatar(ctx, bases)
    (img,), _ = _serve(server, [(aid, coeffs, cam, 48)])
    server.close()
    # direct path on a frame slot the workers do not use
    ctx.set_frame_expression(2, bases, torch.from_numpy(coeffs).cuda())
    ctx.build_occupancy(2, 0.05)
    rgba, _, steps = ctx.render(cam, 48, 48, 2, 2, 0)
    torch.cuda.synchronize()
    assert np.array_equal(img, rgba.cpu().numpy())
    # the oracle on the same blended expression (float32, same order)
    ensure_oracle()
    from oracle.oracle import Oracle
    orc = Oracle(ctx.cfg)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    orc.set_frame(0, synth.blend_expression_f32(av.canonical, av.bases, coeffs))
    orc.set_params(params)
    orc.build_occupancy(0, 0.05)
    o_rgba, _, o_steps, _ = orc.render(cam, 48, 48, 0, 0, 0)
    assert_render_matches(img, steps.cpu().numpy(), o_rgba, o_steps, 1e-5, ctx.cfg.min_transmittance,
                          "served rgba vs oracle")


@pytest.mark.parametrize("graphs", [1, 0])
def test_served_request_conditions_on_its_expression(graphs):
    """16 extra dims: a request's first 16 coefficients become its frame's
    values (captured into the request's graph when graphs are on)."""
    import torch
    av, ctx, bases, params = _avatar_ctx(2, n_frames=3, n_extra=16)
    res = 48
    lips = av.masks == synth.MASK_LIPS
    c = av.canonical[av.faces[lips]].reshape(-1, 3).mean(axis=0) - synth.CENTER
    cam = make_camera(synth.look_at_camera(c / np.linalg.norm(c), 1.0), synth.focal_from_fov(res, 25.0))
    coeffs = np.random.default_rng(6).normal(0.0, 1.0, 50).astype(np.float32)
    server = Server(default_server_config(n_slots=2, hbm_budget_bytes=1 << 30, use_graphs=graphs))
    aid = server.add_avatar(ctx, bases)
    (img,), _ = _serve(server, [(aid, coeffs, cam, res)])
    server.close()
    ctx.set_frame_expression(2, bases, torch.from_numpy(coeffs).cuda())
    ctx.build_occupancy(2, 0.05)
    rgba, _, steps = ctx.render(cam, res, res, 2, 2, 0)
    torch.cuda.synchronize()
    assert np.array_equal(img, rgba.cpu().numpy())
    ensure_oracle()
    from oracle.oracle import Oracle
    orc = Oracle(ctx.cfg)
    orc.set_canonical_mesh(av.canonical, av.faces, av.masks)
    orc.set_frame(0, synth.blend_expression_f32(av.canonical, av.bases, coeffs))
    orc.set_params(params)
    orc.build_occupancy(0, 0.05)
    orc.set_frame_extra_dims(0, coeffs[:16])
    o_rgba, _, o_steps, _ = orc.render(cam, res, res, 0, 0, 0)
    assert_render_matches(img, steps.cpu().numpy(), o_rgba, o_steps, 1e-5, ctx.cfg.min_transmittance,
                          "served rgba vs oracle")
    orc.set_frame_extra_dims(0, np.zeros(16, np.float32))
    assert np.abs(orc.render(cam, res, res, 0, 0, 0)[0] - o_rgba).max() > 1e-3  # the values matter here


@pytest.mark.parametrize("graphs", [1, 0])
def test_same_alone_and_in_mix(graphs):
    avatars = [_avatar_ctx(s, n_frames=4, precision=PRECISION_FP16) for s in (2, 3, 4)]
    server = Server(default_server_config(n_slots=4, hbm_budget_bytes=2 << 30, use_graphs=graphs))
    ids = [server.add_avatar(ctx, bases) for (_, ctx, bases, _) in avatars]
    rng = np.random.default_rng(7)
    c0, cam0 = _request(rng, 96)
    (alone,), _ = _serve(server, [(ids[0], c0, cam0, 96)])
    mix = []
    for k in range(15):
        res = int(rng.choice([32, 64, 96, 128]))
        c, cam = _request(rng, res)
        mix.append((ids[k % 3], c, cam, res))
    mix.insert(7, (ids[0], c0, cam0, 96))
    imgs, results = _serve(server, mix)
    if graphs:
        # each worker instantiates at most one graph per avatar; every other
        # request moved that graph onto its own arguments in place
        updates, builds = server.graph_stats()
        assert updates + builds == len(mix) + 1
        assert builds <= 4 * 3
    server.close()
    assert np.array_equal(imgs[7], alone)
    # the mix really overlapped: some request was admitted before another finished
    spans = sorted((r.t_admit, r.t_done) for r in results.values())
    assert any(b[0] < a[1] for a, b in zip(spans, spans[1:]))
    assert all(np.isfinite(i).all() for i in imgs)


def test_admission_respects_budget():
    import torch
    _, ctx, bases, _ = _avatar_ctx(5, n_frames=4, precision=PRECISION_FP16)
    # a 128x128 request takes the 1M-sample minimum buffer (~70 MB): a 100 MB
    # budget admits one at a time even with 4 idle slots
    server = Server(default_server_config(n_slots=4, hbm_budget_bytes=100 << 20))
    aid = server.add_avatar(ctx, bases)
    rng = np.random.default_rng(9)
    reqs = [(aid,) + _request(rng, 128) + (128,) for _ in range(6)]
    _, results = _serve(server, reqs)
    spans = sorted((r.t_admit, r.t_done) for r in results.values())
    assert all(b[0] >= a[1] for a, b in zip(spans, spans[1:]))
    big = torch.zeros((2048 * 2048, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(InstaError):
        server.submit(aid, np.zeros(50, np.float32), reqs[0][2], 2048, 2048, big)
    server.close()


@pytest.mark.parametrize("graphs", [1, 0])
def test_rounds_beyond_the_queued_ones(graphs):
    """A 1-sample-per-pixel buffer leaves K = capacity / live rays small, so
    most rays need rounds after the two queued with the request (host-checked
    continuation, then the image is rewritten); the image equals the direct
    render, whose buffer is 128 times larger (fewer, longer rounds)."""
    import torch
    av, ctx, bases, _ = _avatar_ctx(6, n_frames=2)
    rng = np.random.default_rng(11)
    coeffs, cam = _request(rng, 512)
    server = Server(default_server_config(n_slots=1, hbm_budget_bytes=1 << 30, samples_per_pixel=1,
                                          use_graphs=graphs))
    aid = server.add_avatar(ctx, bases)
    (img,), results = _serve(server, [(aid, coeffs, cam, 512)])
    server.close()
    ctx.set_frame_expression(1, bases, torch.from_numpy(coeffs).cuda())
    ctx.build_occupancy(1, 0.05)
    rgba, _, _ = ctx.render(cam, 512, 512, 1, 1, 0)
    stats = ctx.render_stats()
    torch.cuda.synchronize()
    assert results[0].n_samples == stats.n_samples
    assert np.array_equal(img, rgba.cpu().numpy())
