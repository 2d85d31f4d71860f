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


def _uvarint(b: bytes, pos: int):
    x, s = 0, 0
    while True:
        c = b[pos]
        pos += 1
        x |= (c & 0x7F) << s
        if c < 0x80:
            return x, pos
        s += 7


def payload_spans(pkt: bytes):
    """Byte ranges of a coded frame's payloads (format.go:376-409 header, the
    golden-override split of codec.go:162-173): [(start, end)] for the frame's
    own payload and, on a key frame with an override, the override's."""
    typ, gov = pkt[0] & 3, (pkt[0] >> 5) & 1
    pos = 6 if typ == 0 else 1
    if not gov:
        return [(pos, len(pkt))]
    kl, p2 = _uvarint(pkt, pos)
    return [(p2, p2 + kl), (p2 + kl, len(pkt))]


def tile_spans(pkt: bytes, payload: int = 0):
    """(offset, length) of each tile stream of one payload (frame.go:197-274:
    qp, tile columns and rows, table block, then length-prefixed tiles)."""
    start, _ = payload_spans(pkt)[payload]
    pos = start + 1
    cols, pos = _uvarint(pkt, pos)
    rows, pos = _uvarint(pkt, pos)
    tlen, pos = _uvarint(pkt, pos)
    pos += tlen
    spans = []
    for _ in range(cols * rows):
        ln, pos = _uvarint(pkt, pos)
        spans.append((pos, ln))
        pos += ln
    return spans


def zero_tile(pkt: bytes, tile: int = 0, payload: int = 0) -> bytes:
    """The packet with one tile's bytes zeroed: the framing still parses, and
    the tile's rANS states start at 0, so its decode underruns."""
    off, ln = tile_spans(pkt, payload)[tile]
    b = bytearray(pkt)
    b[off:off + ln] = bytes(ln)
    return bytes(b)
