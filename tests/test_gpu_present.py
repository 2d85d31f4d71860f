"""Zero-copy presentation: the batch step writes each decoded frame's RGBA
(frame.go:8-67) into caller-owned pitched device surfaces, in turn per
session; compared with the oracle's conversion of the oracle's frame.  A
frame that fails takes no surface; a pitch below 4 * width presents nothing;
the single-frame conversion into a pitched device buffer matches too."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import clip_kw, zero_tile

pytestmark = pytest.mark.gpu


class Surfaces:
    def __init__(self, w, h, count, pad):
        from govid_amd.codec import DeviceBuffer

        self.pitch = 4 * w + pad
        self.w, self.h = w, h
        self.bufs = [DeviceBuffer(bytes(self.pitch * h)) for _ in range(count)]

    @property
    def ptrs(self):
        return [b.ptr for b in self.bufs]

    def read(self, k):
        from govid_amd import native

        out = np.empty((self.h, self.pitch), np.uint8)
        assert native.gpu_lib().ebiv_memcpy_d2h(out.ctypes.data, self.bufs[k].ptr, out.nbytes) == 0
        return out[:, :4 * self.w].reshape(self.h, self.w, 4)

    def close(self):
        for b in self.bufs:
            b.close()


def test_batch_step_presents_into_pitched_surfaces():
    from govid_amd.codec import Batch

    streams = [(list(clip_kw("rects", 98, 66, 9, qp=22, gop=4, tile_cols=2, tile_rows=2).packets), 5, 4, 12),
               (list(clip_kw("parallax", 64, 48, 7, qp=28, gop=7, altref=3).packets), 3, 3, 0),
               (list(clip_kw("synthpan", 37, 21, 6, qp=22, gop=3).packets), 4, 2, 5)]
    streams[0][0][5] = zero_tile(streams[0][0][5])  # fails: takes no surface
    b = Batch()
    sids, surf, want = [], [], []
    for k, (pk, count, window, pad) in enumerate(streams):
        sid = b.open(ring=max(3, window), window=window)
        sids.append(sid)
        ref = O.OracleCodec(ring=max(3, window))
        out = []
        for p in pk:
            try:
                out.append(O.rgba(*ref.decode(p)))
            except O.DecodeError:
                out.append(None)
        want.append(out)
        w, h = (98, 66) if k == 0 else ((64, 48) if k == 1 else (37, 21))
        sf = Surfaces(w, h, count, pad)
        surf.append(sf)
        b.set_surfaces(sid, sf.ptrs, sf.pitch)
        b.submit_many(sid, pk, tags=[(k << 32) | i for i in range(len(pk))])
    next_surface = [0, 0, 0]
    presented = 0
    while b.pending():
        for c in b.step():
            k, i = c.tag >> 32, c.tag & 0xFFFFFFFF
            if want[k][i] is None:
                assert c.status != 0 and c.surface == -1
                continue
            assert c.status == 0 and c.surface == next_surface[k], (k, i, c.surface)
            next_surface[k] = (next_surface[k] + 1) % len(surf[k].bufs)
            got = surf[k].read(c.surface)
            assert np.array_equal(got, want[k][i]), f"stream {k} frame {i}: RGBA differs from the oracle"
            presented += 1
        print(f"  {presented} frames presented", flush=True)
    assert presented == sum(len(s[0]) for s in streams) - 1
    with pytest.raises(Exception):  # fewer surfaces than the window
        b.set_surfaces(sids[0], surf[0].ptrs[:3], surf[0].pitch)
    for sf in surf:
        sf.close()
    b.close()


def test_pitch_below_the_width_presents_nothing():
    from govid_amd.codec import Batch

    pk = list(clip_kw("synthpan", 64, 48, 3, qp=22, gop=3).packets)
    b = Batch()
    sid = b.open(ring=3, window=3)
    sf = Surfaces(63, 48, 3, 0)  # pitch 252 < 4 * 64
    b.set_surfaces(sid, sf.ptrs, sf.pitch)
    b.submit_many(sid, pk)
    done = b.step()
    assert [c.surface for c in done] == [-1, -1, -1] and all(c.status == 0 for c in done)
    sf.close()
    b.close()


def test_single_frame_rgba_into_device_memory():
    from govid_amd.codec import Codec, DeviceBuffer
    from govid_amd import native

    pk = list(clip_kw("rects", 98, 66, 3, qp=22, gop=3).packets)
    gpu, ref = Codec(), O.OracleCodec()
    buf = DeviceBuffer(bytes(98 * 66 * 4 + 16))
    for p in pk:
        f, want = gpu.decode(p), ref.decode(p)
        f.rgba_device(buf.ptr)
        out = np.empty((66, 98, 4), np.uint8)
        assert native.gpu_lib().ebiv_memcpy_d2h(out.ctypes.data, buf.ptr, out.nbytes) == 0
        assert np.array_equal(out, O.rgba(*want))
    buf.close()
    gpu.close()
