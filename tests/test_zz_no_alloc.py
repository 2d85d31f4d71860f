"""ebiv_frame_rgba_noalloc: Frame.RGBA (frame.go:8-67) into host memory
through the library's fixed device staging buffer, compared with
ebiv_frame_rgba (which stages through memory it allocates) and with the
oracle's conversion, on two calls in a row, the second with a smaller frame;
frames above the buffer's 3840x2160 bound are rejected."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from helpers import clip_kw

pytestmark = pytest.mark.gpu


def _rgba_both(codec, frame):
    raw = frame._raw
    a = np.full((frame.height, frame.width, 4), 7, np.uint8)
    b = np.full((frame.height, frame.width, 4), 9, np.uint8)
    L = codec._L
    assert L.ebiv_frame_rgba_noalloc(codec._h, C.byref(raw), a.ctypes.data) == 0, L.ebiv_codec_last_error(codec._h)
    assert L.ebiv_frame_rgba(codec._h, C.byref(raw), b.ctypes.data, 0) == 0
    return a, b


def test_noalloc_rgba_matches_allocating_entry_point():
    from govid_amd.codec import Codec
    from govid_amd.container import Reader

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    big = Reader.open(os.path.join(root, "bench_data", "parallax_2160p.ebiv")).packets()[:2]
    small = clip_kw("rects", 320, 180, 3, qp=22, gop=3).packets
    c1, c2 = Codec(ring=3), Codec(ring=3)
    ref1, ref2 = O.OracleCodec(), O.OracleCodec()
    for p in big:
        f1, want1 = c1.decode(p), ref1.decode(p)
    print(f"first call: {f1.width}x{f1.height}", flush=True)
    a, b = _rgba_both(c1, f1)
    assert (f1.width, f1.height) == (3840, 2160)
    assert np.array_equal(a, b) and np.array_equal(a, O.rgba(*want1))
    for p in small:
        f2, want2 = c2.decode(p), ref2.decode(p)
    print(f"second call: {f2.width}x{f2.height}", flush=True)
    a2, b2 = _rgba_both(c2, f2)
    assert np.array_equal(a2, b2) and np.array_equal(a2, O.rgba(*want2))
    # the first codec's frame again, after the smaller call reused the buffer
    a3, _ = _rgba_both(c1, f1)
    assert np.array_equal(a3, a)
    c1.close()
    c2.close()


def test_noalloc_rgba_rejects_frames_above_its_bound():
    from govid_amd import capi
    from govid_amd.codec import Codec

    w, h = 4096, 2304
    n = w * h + 2 * (w // 2) * (h // 2)
    key = bytes([0x00, w & 255, w >> 8, h & 255, h >> 8, 0]) + bytes(n)
    c = Codec(ring=2)
    f = c.decode(key)
    out = np.zeros((h, w, 4), np.uint8)
    rc = c._L.ebiv_frame_rgba_noalloc(c._h, C.byref(f._raw), out.ctypes.data)
    assert rc == capi.ERR_ARGUMENT
    assert b"too large" in c._L.ebiv_codec_last_error(c._h)
    c.close()
