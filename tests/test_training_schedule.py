"""Host schedule of the subset reload (insta_amd.training): when reloads
happen (src/core.cu:684, testbed_nerf.cu:2976-2978) and what a subset holds
(nerf_loader.cu:461-495).  Which images a subset holds is not the
reference's (std::mt19937 + std::shuffle): parity unpinned there; the
counts, ranges and the persistence of the generator are checked."""
import numpy as np
import pytest

from insta_amd.training import IMAGES_TO_LOAD, RELOAD_EVERY, SubsetSampler, reload_due


def test_reload_steps():
    due = [s for s in range(0, 6001) if reload_due(s)]
    assert due == [1500, 3000, 4500, 6000]
    assert not reload_due(0) and RELOAD_EVERY == 1500 and IMAGES_TO_LOAD == 1700
    assert [s for s in range(1, 13) if reload_due(s, 5)] == [5, 10]


def test_subset_takes_min_n_images_per_file():
    s = SubsetSampler([10, 3, 7], images_to_load=5, seed=42)
    ids = s.draw()
    assert len(ids) == 5 + 3 + 5
    a, b, c = ids[:5], ids[5:8], ids[8:]
    assert set(a) <= set(range(0, 10)) and len(set(a)) == 5
    assert sorted(b) == [10, 11, 12]                  # fewer than images_to_load: all of them
    assert set(c) <= set(range(13, 20)) and len(set(c)) == 5


def test_generator_persists_across_reloads():
    s = SubsetSampler([200], images_to_load=20, seed=42)
    draws = [tuple(s.draw()) for _ in range(5)]
    assert len(set(draws)) == 5                       # each reload a new subset
    # a new sampler with the same seed replays the same sequence
    t = SubsetSampler([200], images_to_load=20, seed=42)
    assert [tuple(t.draw()) for _ in range(5)] == draws
    # over many reloads every image is drawn, about equally often
    counts = np.zeros(200)
    for _ in range(400):
        counts[s.draw()] += 1
    assert counts.min() > 0 and abs(counts.mean() - 40) < 1e-9 and counts.std() < 12


def test_unshuffled_subset_is_the_first_ids():
    s = SubsetSampler([6, 4], images_to_load=3, shuffle=False)
    assert s.draw() == [0, 1, 2, 6, 7, 8] == s.draw()


def test_bad_sizes_rejected():
    with pytest.raises(ValueError):
        SubsetSampler([3], images_to_load=0)
