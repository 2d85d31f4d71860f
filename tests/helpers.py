"""Shared test helpers: clip encoding with caching, plane diffs (harness_test.go:46-99)."""
from __future__ import annotations

import functools

import numpy as np

from govid_amd import clips, encoder


@functools.lru_cache(maxsize=64)
def clip(workload: str, w: int, h: int, frames: int, seed: int = 0, **kw):
    kw_t = dict(kw)
    frames_l = list(clips.generate(workload, w, h, frames, clips.stream_seed(workload, seed)))
    return encoder.encode_frames(frames_l, w, h, **kw_t)


def clip_kw(workload, w, h, frames, seed=0, **kw):
    return clip(workload, w, h, frames, seed, **kw)


def plane_diff(got: np.ndarray, want: np.ndarray):
    """comparePlanes: wrong count, max error, first error (x, y)."""
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    wrong = int((d != 0).sum())
    if wrong == 0:
        return 0, 0, None
    ys, xs = np.nonzero(d)
    return wrong, int(d.max()), (int(xs[0]), int(ys[0]))


def assert_planes_equal(got, want, label=""):
    for name, a, b in zip(("Y", "Cb", "Cr"), got, want):
        assert a.shape == b.shape, f"{label} {name}: shape {a.shape} != {b.shape}"
        wrong, mx, first = plane_diff(a, b)
        assert wrong == 0, f"{label} plane {name}: {wrong}/{a.size} samples wrong, max error {mx}, first at {first}"
