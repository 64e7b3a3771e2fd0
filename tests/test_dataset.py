"""Dataset ingestion (insta_amd.dataset) on a capture written here: frames
sorted by mesh path, OBJ meshes with the +0.5 offset, FLAME parameter files,
camera matrices with the scene offset, focal lengths and principal points
from the reference's keys, images decoded to RGBA floats."""
import json
import math

import numpy as np
import pytest

from insta_amd import dataset, synth


def _write_obj(path, verts, faces, style="plain"):
    lines = [f"v {x:.9g} {y:.9g} {z:.9g}" for x, y, z in verts]
    for a, b, c in faces:
        if style == "slash":
            lines.append(f"f {a + 1}/{a + 1}/{a + 1} {b + 1}//{b + 1} {c + 1}/{c + 1}")
        else:
            lines.append(f"f {a + 1} {b + 1} {c + 1}")
    path.write_text("\n".join(lines) + "\n")


@pytest.fixture
def capture(tmp_path):
    from PIL import Image
    av = synth.small_sphere_avatar(seed=4, n_frames=3)
    canon = av.canonical - 0.5            # the loader adds 0.5
    _write_obj(tmp_path / "canonical.obj", canon, av.faces)
    (tmp_path / "meshes").mkdir()
    (tmp_path / "exp").mkdir()
    (tmp_path / "images").mkdir()
    rng = np.random.default_rng(0)
    frames = []
    imgs = []
    for i, v in enumerate(av.frames):
        _write_obj(tmp_path / "meshes" / f"{i:05d}.obj", v - 0.5, av.faces, style="slash" if i == 1 else "plain")
        (tmp_path / "exp" / f"{i:05d}.txt").write_text("\n".join(f"{x:.9g}" for x in rng.normal(size=100)) + "\n")
        img = rng.integers(0, 256, (24, 32, 4), dtype=np.uint8)
        Image.fromarray(img, "RGBA").save(tmp_path / "images" / f"{i:05d}.png")
        imgs.append(img)
        m = np.eye(4)
        m[:3, 3] = rng.normal(size=3)
        frames.append({"file_path": f"images/{i:05d}", "mesh_path": f"meshes/{i:05d}.obj",
                       "exp_path": f"exp/{i:05d}.txt", "transform_matrix": m.tolist()})
    frames[2]["camera_angle_x"] = 0.7          # frame-level focal (radians)
    frames[2]["cx"], frames[2]["w"], frames[2]["cy"], frames[2]["h"] = 8.0, 32, 6.0, 24
    doc = {"fl_x": 40.0, "fl_y": 41.0, "cx": 16.5, "cy": 11.0, "w": 32, "h": 24,
           "frames": [frames[2], frames[0], frames[1]]}           # not in mesh-path order
    (tmp_path / "transforms_train.json").write_text(json.dumps(doc))
    return tmp_path, av, doc, imgs


def test_capture_loads_with_the_reference_conventions(capture):
    path, av, doc, imgs = capture
    ds = dataset.load_transforms(path / "transforms_train.json")
    # written as v - 0.5, read back + 0.5 in float32: equal to rounding
    assert np.allclose(ds.canonical, av.canonical, atol=1e-6) and np.array_equal(ds.faces, av.faces)
    assert (ds.masks == synth.MASK_TORSO).all()
    assert len(ds.frames) == 3 and ds.resolution == (32, 24)
    for i in range(3):                     # sorted by mesh path: image i is mesh i
        assert np.allclose(ds.frames[i], av.frames[i], atol=1e-6)
        assert ds.image_paths[i].name == f"{i:05d}.png"
        assert ds.expressions[i].shape == (75,)
    by_mesh = sorted(doc["frames"], key=lambda f: f["mesh_path"])
    for cam, fr in zip(ds.cameras, by_mesh):
        m = np.asarray(fr["transform_matrix"])[:3, :4].copy()
        m[:, 3] += 0.5
        assert np.allclose(np.frombuffer(bytes(cam.matrix), np.float32).reshape(3, 4), m, atol=1e-6)
    c0, c2 = ds.cameras[0], ds.cameras[2]
    assert (c0.focal_px[0], c0.focal_px[1]) == (40.0, 41.0)
    assert c0.screen_center[0] == pytest.approx(16.5 / 32) and c0.screen_center[1] == pytest.approx(11.0 / 24)
    assert c2.focal_px[0] == pytest.approx(0.5 * 32 / math.tan(0.35), rel=1e-6) == c2.focal_px[1]
    assert (c2.screen_center[0], c2.screen_center[1]) == (pytest.approx(0.25), pytest.approx(0.25))
    got = ds.images()
    assert np.array_equal(got, np.stack(imgs).astype(np.float32) / 255.0)


def test_n_frames_offset_and_aabb(capture):
    path, _, doc, _ = capture
    doc = dict(doc, n_frames=2, offset=[0.1, 0.2, 0.3])
    (path / "t2.json").write_text(json.dumps(doc))
    ds = dataset.load_transforms(path / "t2.json")
    assert len(ds.cameras) == 2 and ds.offset == (0.1, 0.2, 0.3)
    m = np.frombuffer(bytes(ds.cameras[0].matrix), np.float32).reshape(3, 4)
    t = np.asarray(sorted(doc["frames"], key=lambda f: f["mesh_path"])[0]["transform_matrix"])[:3, 3]
    assert np.allclose(m[:, 3], t + [0.1, 0.2, 0.3], atol=1e-6)
    doc3 = dict(doc, aabb=[[-1.0, -2.0, 0.0], [1.0, 2.0, 1.0]])
    (path / "t3.json").write_text(json.dumps(doc3))
    ds3 = dataset.load_transforms(path / "t3.json")
    assert ds3.scale == pytest.approx(0.25) and np.allclose(ds3.offset, [0.5, 0.5, 0.375])


def _add_supervision(path, doc, rng, depth_files=("png", "bin", None)):
    """Segmentation masks for every frame and depth files per frame (16-bit
    PNG, .bin of uint32, or none); returns the written ids and raw depths by
    mesh order."""
    from PIL import Image
    (path / "seg").mkdir(exist_ok=True)
    (path / "depth").mkdir(exist_ok=True)
    ids, raws = [], []
    by_mesh = sorted(doc["frames"], key=lambda f: f["mesh_path"])
    for i, fr in enumerate(by_mesh):
        seg = rng.integers(0, 256, (24, 32, 4), dtype=np.uint8)
        Image.fromarray(seg, "RGBA").save(path / "seg" / f"{i:05d}.png")
        fr["seg_mask_path"] = f"seg/{i:05d}.png"
        ids.append(seg[..., 0])
        raw = rng.integers(0, 65536, (24, 32)).astype(np.uint16)
        kind = depth_files[i]
        if kind == "png":
            Image.fromarray(raw).save(path / "depth" / f"{i:05d}.png")
            fr["depth_path"] = f"depth/{i:05d}.png"
        elif kind == "bin":
            # the upper 16 bits are dropped by the reader (uint32 -> uint16)
            (raw.astype("<u4") | (rng.integers(0, 3, raw.shape).astype("<u4") << 16)).tofile(
                path / "depth" / f"{i:05d}.bin")
            fr["depth_path"] = f"depth/{i:05d}.bin"
        else:
            raw = np.zeros_like(raw)
        raws.append(raw)
    return ids, raws


def test_segmentation_masks_and_depth_images(capture):
    path, _, doc, _ = capture
    rng = np.random.default_rng(7)
    doc = json.loads(json.dumps(doc))
    ids, raws = _add_supervision(path, doc, rng)
    doc["integer_depth_scale"] = 0.001
    doc["scale"] = 1.5
    (path / "sup.json").write_text(json.dumps(doc))
    ds = dataset.load_transforms(path / "sup.json")
    seg = ds.seg_masks()
    assert seg.dtype == np.uint8 and np.array_equal(seg, np.stack(ids))
    depth = ds.depth_images()
    expect = np.stack(raws).astype(np.float32) * np.float32(0.001 * 1.5)
    assert depth.dtype == np.float32 and np.array_equal(depth, expect)
    assert (depth[2] == 0).all() and (depth[0] > 0).mean() > 0.99
    assert np.array_equal(ds.seg_masks([2, 0]), seg[[2, 0]])
    assert np.array_equal(ds.depth_images([1]), depth[[1]])
    # no integer_depth_scale: no depth at all, as the reference skips the files
    del doc["integer_depth_scale"]
    (path / "nodepth.json").write_text(json.dumps(doc))
    assert dataset.load_transforms(path / "nodepth.json").depth_images() is None
    # masks for only some images of a set are refused (the device takes all or none)
    for fr in doc["frames"]:
        if fr["mesh_path"].endswith("00001.obj"):
            del fr["seg_mask_path"]
    (path / "someseg.json").write_text(json.dumps(doc))
    ds2 = dataset.load_transforms(path / "someseg.json")
    assert ds2.seg_masks([0, 2]) is not None
    with pytest.raises(ValueError):
        ds2.seg_masks()


def test_8_bit_depth_images_scale_as_stbi_load_16(tmp_path):
    from PIL import Image
    a = np.arange(12, dtype=np.uint8).reshape(3, 4) * 20
    Image.fromarray(a, "L").save(tmp_path / "d.png")
    assert np.array_equal(dataset.read_depth(tmp_path / "d.png", 4, 3), a.astype(np.uint16) * 257)
    Image.fromarray(np.zeros((3, 4, 3), np.uint8), "RGB").save(tmp_path / "rgb.png")
    with pytest.raises(ValueError):
        dataset.read_depth(tmp_path / "rgb.png", 4, 3)


def test_obj_polygons_and_negative_indices(tmp_path):
    p = tmp_path / "q.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\nf -4 -2 -1\n")
    v, f = dataset.read_obj(p)
    assert np.allclose(v, np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]) + 0.5)
    assert f.tolist() == [[0, 1, 2], [0, 2, 3], [0, 2, 3]]
    p.write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(ValueError):
        dataset.read_obj(p)


def test_inconsistent_captures_are_rejected(capture):
    path, av, doc, _ = capture
    _write_obj(path / "meshes" / "00001.obj", av.frames[1][:-1] - 0.5, av.faces[:-1])
    with pytest.raises(ValueError):
        dataset.load_transforms(path / "transforms_train.json")
    doc2 = {k: v for k, v in doc.items() if k not in ("fl_x", "fl_y")}
    doc2["frames"] = [f for f in doc["frames"] if "camera_angle_x" not in f]
    (path / "nofocal.json").write_text(json.dumps(doc2))
    with pytest.raises(ValueError):
        dataset.load_transforms(path / "nofocal.json")


def _write_mask_objs(folder, verts, faces, members, colours=None):
    """mesh_masks/*.obj: one file per region, red (1, 0, 0) on its member
    vertices, grey elsewhere; `colours` overrides (region, vertex) -> rgb."""
    folder.mkdir(parents=True, exist_ok=True)
    for name, _ in dataset.MESH_MASK_REGIONS:
        lines = ["# mask"]
        for i, v in enumerate(verts):
            rgb = (1.0, 0.0, 0.0) if i in members.get(name, ()) else (0.588235, 0.588235, 0.588235)
            rgb = (colours or {}).get((name, i), rgb)
            lines.append("v %f %f %f" % tuple(v) + ("" if rgb is None else " %f %f %f" % tuple(rgb)))
        lines += ["f %d %d %d" % tuple(f + 1) for f in faces]
        (folder / f"{name}.obj").write_text("\n".join(lines) + "\n")


def _triangle_to_mask_loops(faces, members):
    """Masking::triangle_to_mask (src/masking.cpp:120-148) as its loops read:
    regions in FlameMasks order, the first member vertex found in a face sets
    the face's id, a later region overwrites an earlier one."""
    out = [15] * len(faces)
    for fi, face in enumerate(faces):
        for name, rid in dataset.MESH_MASK_REGIONS:
            for v in members.get(name, []):
                if sum(int(face[k] == v) for k in range(3)) > 0:
                    out[fi] = rid
                    break
    return out


def test_region_masks_follow_the_reference_rule(tmp_path):
    verts = np.random.default_rng(0).uniform(-0.1, 0.1, (9, 3))
    faces = np.array([[0, 1, 2], [2, 3, 4], [5, 6, 7], [6, 7, 8], [1, 3, 5]])
    members = {"lips": [0], "face": [0, 3], "right_ear": [4], "boundary": [6], "eye_region": [8]}
    _write_mask_objs(tmp_path / "mesh_masks", verts, faces, members)
    got = dataset.region_masks(faces, len(verts), tmp_path / "mesh_masks")
    # lips (7) over face (2); right ear (10) over face; boundary only; eye
    # region (1) over boundary (0); face (2) through vertex 3
    assert got.tolist() == [7, 10, 0, 1, 2]
    # colour threshold of load_mesh_color_mask: r >= 1, g == b == 0
    colours = {("lips", 0): (0.99, 0.0, 0.0), ("face", 0): (1.0, 0.001, 0.0), ("face", 3): (1.5, 0.0, 0.0),
               ("right_ear", 4): None, ("boundary", 6): (1.0, 0.0, 0.0)}
    _write_mask_objs(tmp_path / "m2", verts, faces, members, colours)
    got = dataset.region_masks(faces, len(verts), tmp_path / "m2")
    assert got.tolist() == [15, 2, 0, 1, 2]


def test_region_masks_match_the_reference_loops_on_a_random_mesh(tmp_path):
    rng = np.random.default_rng(3)
    nv = 300
    verts = rng.uniform(-0.1, 0.1, (nv, 3))
    faces = np.stack([rng.choice(nv, 3, replace=False) for _ in range(500)])
    members = {name: sorted(rng.choice(nv, int(rng.integers(5, 60)), replace=False).tolist())
               for name, _ in dataset.MESH_MASK_REGIONS}
    _write_mask_objs(tmp_path / "mesh_masks", verts, faces, members)
    got = dataset.region_masks(faces, nv, tmp_path / "mesh_masks")
    assert got.tolist() == _triangle_to_mask_loops(faces, members)
    assert len(set(got.tolist())) >= 5


def test_capture_takes_its_masks_from_mesh_masks(capture, tmp_path):
    path = capture[0] / "transforms_train.json"
    ds0 = dataset.load_transforms(path)
    members = {"lips": list(range(0, len(ds0.canonical), 7))}
    _write_mask_objs(tmp_path / "mesh_masks", ds0.canonical, ds0.faces, members)
    ds = dataset.load_transforms(path, mesh_masks_dir=tmp_path / "mesh_masks")
    lips_faces = np.isin(ds0.faces, members["lips"]).any(axis=1)
    assert np.array_equal(ds.masks == synth.MASK_LIPS, lips_faces)
    assert np.all(ds.masks[~lips_faces] == synth.MASK_TORSO)
    with pytest.raises(ValueError):
        _write_mask_objs(tmp_path / "bad", ds0.canonical[:-1], ds0.faces[:1], members)
        dataset.load_transforms(path, mesh_masks_dir=tmp_path / "bad")
