"""insta_load_config_json (Testbed::load_network_config with
merge_parent_network_config, instant-ngp/src/testbed.cu:77-152, read as
reset_network / NerfNetwork / rta::Core read it) on configuration files
shaped like the reference's configs/nerf/main.json and insta.json.  Host-only
entry point: runs without a GPU."""
import json
import math

import pytest

from insta_amd.capi import InstaError, load_config_json

# the keys and values of a main.json-style network configuration
MAIN = {
    "loss": {"otype": "Huber"},
    "nerf_network": "concatenation",
    "max_steps": 33000, "max_cached_bvh": 4000, "max_images_gpu": 1700, "use_dataset_cache": True,
    "use_eyes": False, "optimize_latent_code": False, "geo_params": 0, "exp_params": 16,
    "seg_weight": 40, "depth_weight": 1.25,
    "optimizer": {"otype": "Ema", "decay": 0.95,
                  "nested": {"otype": "Adam", "learning_rate": 0.00175, "beta1": 0.9, "beta2": 0.99,
                             "epsilon": 1e-15, "l2_reg": 1e-6}},
    "encoding": {"otype": "HashGrid", "n_levels": 16, "n_features_per_level": 8, "log2_hashmap_size": 18,
                 "base_resolution": 16},
    "network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None", "n_neurons": 64,
                "n_hidden_layers": 2, "n_output_dims": 16},
    "dir_encoding": {"otype": "Composite", "nested": [{"n_dims_to_encode": 3, "otype": "SphericalHarmonics",
                                                       "degree": 4}, {"otype": "Identity"}]},
    "rgb_network": {"otype": "FullyFusedMLP", "activation": "ReLU", "output_activation": "None", "n_neurons": 64,
                    "n_hidden_layers": 3},
}


def _write(path, doc, comment=True):
    text = json.dumps(doc, indent=1)
    if comment:  # the reference parses with comments allowed
        text = "// network configuration\n/* block\n comment */" + text
    path.write_text(text)
    return path


def test_main_config_fields(tmp_path):
    cfg, run = load_config_json(_write(tmp_path / "main.json", MAIN))
    assert (cfg.n_levels, cfg.n_features_per_level, cfg.log2_hashmap_size, cfg.base_resolution) == (16, 8, 18, 16)
    # per_level_scale from the finest resolution 2048 (testbed.cu:2254-2259)
    assert math.isclose(cfg.per_level_scale, math.exp(math.log(2048 / 16) / 15), rel_tol=1e-6)
    assert (cfg.n_neurons, cfg.density_hidden_layers, cfg.density_out, cfg.rgb_hidden_layers) == (64, 2, 16, 3)
    assert cfg.sh_degree == 4 and cfg.n_extra_dims == 16 and cfg.disable_interp == 0
    assert math.isclose(run.ema_decay, 0.95, rel_tol=1e-6) and math.isclose(run.learning_rate, 0.00175, rel_tol=1e-6)
    assert (run.beta1, run.beta2) == pytest.approx((0.9, 0.99)) and run.epsilon == pytest.approx(1e-15)
    assert run.l2_reg == pytest.approx(1e-6) and run.exp_params == 16 and run.max_training_steps == 33000
    assert run.depth_weight == pytest.approx(1.25) and run.seg_weight == pytest.approx(40.0)
    assert run.max_cached_bvh == 4000 and run.max_images_gpu == 1700


def test_parent_chain_merge_patch(tmp_path):
    _write(tmp_path / "main.json", MAIN)
    (tmp_path / "sub").mkdir()
    # insta.json style child, one directory down with a relative parent
    child = {"parent": "../main.json", "max_steps": 20000, "disable_interp": True,
             "optimizer": {"nested": {"learning_rate": 0.01, "l2_reg": None}},
             "encoding": {"n_features_per_level": 2, "log2_hashmap_size": 19, "per_level_scale": 1.5}}
    cfg, run = load_config_json(_write(tmp_path / "sub" / "insta.json", child))
    assert run.max_training_steps == 20000 and cfg.disable_interp == 1
    # nested objects merge key by key; null removes the key (default 0)
    assert run.learning_rate == pytest.approx(0.01) and run.beta2 == pytest.approx(0.99)
    assert run.l2_reg == 0.0 and run.ema_decay == pytest.approx(0.95)
    assert (cfg.n_levels, cfg.n_features_per_level, cfg.log2_hashmap_size) == (16, 2, 19)
    assert cfg.per_level_scale == pytest.approx(1.5)
    # grandchild on top of the child
    grand = {"parent": "sub/insta.json", "exp_params": 0, "network": {"n_hidden_layers": 1}}
    cfg, run = load_config_json(_write(tmp_path / "grand.json", grand))
    assert cfg.n_extra_dims == 0 and cfg.density_hidden_layers == 1 and run.max_training_steps == 20000


def test_n_features_and_default_base_resolution(tmp_path):
    doc = {"encoding": {"otype": "HashGrid", "n_features": 32, "n_features_per_level": 2, "log2_hashmap_size": 15}}
    cfg, _ = load_config_json(_write(tmp_path / "a.json", doc, comment=False))
    # n_levels = n_features / F; base resolution 2^(log2_T / 3) (testbed.cu:2228-2244)
    assert cfg.n_levels == 16 and cfg.base_resolution == 1 << 5


@pytest.mark.parametrize("patch", [
    {"encoding": {"otype": "Frequency"}},
    {"loss": {"otype": "L2"}},
    {"network": {"activation": "Sine"}},
    {"optimizer": {"otype": "Shampoo"}},
    {"dir_encoding": {"otype": "OneBlob"}},
])
def test_unsupported_choices_are_rejected(tmp_path, patch):
    _write(tmp_path / "main.json", MAIN)
    with pytest.raises(InstaError) as e:
        load_config_json(_write(tmp_path / "bad.json", {"parent": "main.json", **patch}))
    assert e.value.status == 5


def test_malformed_and_missing_files(tmp_path):
    (tmp_path / "broken.json").write_text('{"encoding": {"n_levels": 16,}')
    with pytest.raises(InstaError):
        load_config_json(tmp_path / "broken.json")
    with pytest.raises(InstaError):
        load_config_json(tmp_path / "absent.json")
    with pytest.raises(InstaError):
        load_config_json(_write(tmp_path / "orphan.json", {"parent": "nowhere.json"}))
    # a parent cycle ends at the depth limit instead of recursing forever
    _write(tmp_path / "a.json", {"parent": "b.json"})
    _write(tmp_path / "b.json", {"parent": "a.json"})
    with pytest.raises(InstaError):
        load_config_json(tmp_path / "a.json")
