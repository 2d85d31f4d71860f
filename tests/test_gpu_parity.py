"""GPU decode vs the oracle, sample for sample, through the C ABI."""
import numpy as np
import pytest

import oracle as O
from helpers import assert_planes_equal, clip_kw

pytestmark = pytest.mark.gpu


def _run(clip, ring=3, check_rgba=False):
    from govid_amd.codec import Codec

    gpu = Codec(ring=ring)
    ref = O.OracleCodec(ring=ring)
    for i, pkt in enumerate(clip.packets):
        print(f"  frame {i} ({len(pkt)} B)", flush=True)
        f = gpu.decode(pkt)
        want = ref.decode(pkt)
        assert_planes_equal(f.planes(), want, f"frame {i}")
        if check_rgba:
            assert np.array_equal(f.rgba(), O.rgba(*want)), f"frame {i} rgba"
    gpu.close()


@pytest.mark.parametrize("case", [
    dict(workload="synthframe", w=64, h=48, frames=3, qp=20, gop=1),
    dict(workload="synthpan", w=64, h=48, frames=6, qp=22, gop=6),
    dict(workload="rects", w=64, h=48, frames=6, qp=22, gop=3, tile_cols=2, tile_rows=2),
    dict(workload="synthpan", w=37, h=21, frames=5, qp=22, gop=5),
    dict(workload="parallax", w=64, h=48, frames=9, qp=28, gop=9, altref=4),
])
def test_small_clips(case):
    c = dict(case)
    wl, w, h, n = c.pop("workload"), c.pop("w"), c.pop("h"), c.pop("frames")
    print(f"case {wl} {w}x{h} x{n} {c}", flush=True)
    _run(clip_kw(wl, w, h, n, threads=4, **c), check_rgba=True)


def test_bench_clip_1080p_all_frames():
    """The benchmark's 1080p 4x3-tile clip (120 frames, two key frames, the
    flash frames) decodes bit-exactly, every frame."""
    import os

    from govid_amd.container import Reader

    path = os.path.join(os.path.dirname(os.path.dirname(__file__)), "bench_data", "parallax_1080p.ebiv")
    rd = Reader.open(path)

    class Clip:
        packets = rd.packets()

    from govid_amd.codec import Codec

    gpu = Codec(ring=3)
    ref = O.OracleCodec(ring=3, threads=8)
    for i, pkt in enumerate(Clip.packets):
        got = gpu.decode(pkt).planes()
        want = ref.decode(pkt)
        assert_planes_equal(got, want, f"frame {i}")
    gpu.close()
