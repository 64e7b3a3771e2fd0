"""Host side of the reference's training run over a capture larger than the
device holds: a random subset of the training images is loaded, trained on,
and replaced by a fresh subset every `reload_every` training steps.

Follows the reference:
  * rta::Core::reload_training_data (src/core.cu:663-706), called after
    every optimizer step (testbed_nerf.cu:2976-2978): when the training step
    is a positive multiple of 1500 the dataset, its meshes and their frames
    are loaded again;
  * load_nerf's subset (instant-ngp/src/nerf_loader.cu:461-495): per
    transforms file, the frame ids 0..n-1 are shuffled by one generator that
    lives across reloads (a global std::mt19937 seeded 42) and the first
    min(n, images_to_load) are kept; images_to_load is "max_images_gpu"
    (configs/nerf/main.json: 1700; src/core.cu:205).

The generator is numpy's PCG64 seeded 42, not std::mt19937 + std::shuffle
(whose permutation is library-defined): which images a subset holds is not
the reference's, only how many, how they are drawn and when.

Each subset is uploaded as frames 0..k-1 of the context (posed meshes), the
training set (images, cameras, those frames) and, where the capture has
them, its segmentation masks and depth images; training then runs
insta_train_loop (density-grid updates on the reference schedule into one
occupancy slot that every ray marches) up to the next reload.
"""
from __future__ import annotations

import numpy as np

RELOAD_EVERY = 1500          # src/core.cu:684
IMAGES_TO_LOAD = 1700        # DatasetSettings::images_to_load (nerf_loader.h:219)


def reload_due(training_step: int, every: int = RELOAD_EVERY) -> bool:
    """True after the optimizer step that brings the step count to a
    positive multiple of `every` (fast_mod(m_training_step, 1500) == 0)."""
    return training_step > 0 and training_step % every == 0


class SubsetSampler:
    """Training-image subsets of a capture whose transforms files hold
    `file_counts` images each (in the order the dataset lists them)."""

    def __init__(self, file_counts, images_to_load: int = IMAGES_TO_LOAD, shuffle: bool = True,
                 seed: int = 42):
        if images_to_load <= 0:
            raise ValueError("images_to_load must be positive")
        self.file_counts = [int(c) for c in file_counts]
        self.images_to_load = int(images_to_load)
        self.shuffle = shuffle
        self.rng = np.random.Generator(np.random.PCG64(seed))

    def draw(self) -> list[int]:
        """Dataset indices of the next subset: per file, the first
        min(n, images_to_load) of a fresh shuffle of its ids."""
        out, base = [], 0
        for n in self.file_counts:
            ids = np.arange(n)
            if self.shuffle:
                ids = self.rng.permutation(n)
            out += [base + int(i) for i in ids[:min(n, self.images_to_load)]]
            base += n
        return out


class Trainer:
    """Trains `ctx` on `ds` (an insta_amd.dataset.Dataset) subset by subset.
    The context must hold at least as many frames as a subset has images.
    The caller's `tc` is used for every step; its seed moves on at each
    reload.  frame_occupancy_threshold: also build each frame's mesh
    occupancy grid at this distance (None: only the density slot)."""

    def __init__(self, ctx, ds, tc, images_to_load: int = IMAGES_TO_LOAD, reload_every: int = RELOAD_EVERY,
                 density_slot: int = -1, density_decay: float = 0.95, density_seed: int = 1337,
                 subset_seed: int = 42, shuffle: bool = True, frame_occupancy_threshold: float | None = None):
        self.ctx, self.ds, self.tc = ctx, ds, tc
        self.reload_every = int(reload_every)
        if self.reload_every <= 0:
            raise ValueError("reload_every must be positive")
        self.density_slot, self.density_decay, self.density_seed = density_slot, density_decay, density_seed
        self.frame_threshold = frame_occupancy_threshold
        self.sampler = SubsetSampler(ds.file_counts or [len(ds.cameras)], images_to_load, shuffle, subset_seed)
        self.subset: list[int] = []
        self.loads = 0
        self.step_count = 0   # the context's training step after the last chunk
        self.last = None

    def load_subset(self) -> list[int]:
        """Draw the next subset and upload it: posed meshes into frames
        0..k-1, images, cameras, masks and depth images."""
        import torch
        ids = self.sampler.draw()
        k = len(ids)
        if k > self.ctx.cfg.max_frames:
            raise ValueError(f"a subset of {k} images needs max_frames >= {k}")
        ctx, ds = self.ctx, self.ds
        ctx.set_frames(0, torch.from_numpy(np.stack([ds.frames[i] for i in ids]).astype(np.float32)).cuda())
        if self.frame_threshold is not None:
            # per-frame mesh occupancy (and the deformation's candidate lists);
            # the rays themselves march the density slot
            for f in range(k):
                ctx.build_occupancy(f, self.frame_threshold)
        ctx.set_training_set(torch.from_numpy(ds.images(ids)).cuda(), [ds.cameras[i] for i in ids], list(range(k)))
        seg, depth = ds.seg_masks(ids), ds.depth_images(ids)
        if seg is not None or depth is not None:
            ctx.set_training_supervision(None if seg is None else torch.from_numpy(seg).cuda(),
                                         None if depth is None else torch.from_numpy(depth).cuda())
        # a fresh pixel stream per subset (the training set restarts its ray count)
        self.tc.seed = (self.tc.seed + 0x9E3779B9) & 0xFFFFFFFF if self.loads else self.tc.seed
        self.subset = ids
        self.loads += 1
        return ids

    def train(self, n_steps: int):
        """n_steps optimizer steps; a new subset is loaded before the first
        and after every step that makes the count a multiple of reload_every."""
        if not self.subset:
            self.load_subset()
        left = int(n_steps)
        while left > 0:
            chunk = min(left, self.reload_every - self.step_count % self.reload_every)
            self.last = self.ctx.train_loop(self.tc, chunk, density_slot=self.density_slot,
                                            n_frames=len(self.subset), decay=self.density_decay,
                                            seed=self.density_seed)
            self.step_count = int(self.last.step)   # the context's training step
            left -= chunk
            if reload_due(self.step_count, self.reload_every):
                self.load_subset()
        return self.last
