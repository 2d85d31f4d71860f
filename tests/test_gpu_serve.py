"""The serving loop on the GPU: a seeded Poisson mix of plays, seeks and loop
restarts over streams of different sizes, admitted under an HBM budget that
forces requests to wait.  Every frame equals the oracle's straight-through
decode, and a request's frames are the same served alone as in the mix."""
import zlib

import pytest

import oracle as O
from govid_amd.serve import Server, Stream, poisson_requests
from helpers import clip_kw

pytestmark = pytest.mark.gpu


def _streams():
    out = []
    for wl, w, h, n, kw in (("rects", 320, 180, 30, dict(qp=22, gop=10)),
                            ("parallax", 128, 64, 20, dict(qp=22, gop=10, tile_cols=2, tile_rows=2)),
                            ("noisepan", 128, 96, 12, dict(qp=12, gop=4, tile_cols=2, tile_rows=2)),
                            ("parallax", 64, 48, 9, dict(qp=28, gop=9, altref=4))):
        c = clip_kw(wl, w, h, n, threads=4, **kw)
        out.append(Stream(f"{wl}{w}x{h}", w, h, c.packets))
    return out


def _crc(planes):
    return tuple(zlib.crc32(p.tobytes()) for p in planes)


def _serve(reqs, budget, window=4):
    from govid_amd.codec import Batch, session_bytes

    got = {}

    def on_frame(fd):
        assert fd.status == 0, (fd.rid, fd.seq)
        got.setdefault(fd.rid, []).append((fd.seq, fd.index, _crc(fd.frame.planes())))

    b = Batch()
    srv = Server(b, budget, window=window, ring=window, on_frame=on_frame)
    stats = srv.run(reqs, max_seconds=120)
    b.close()
    return got, stats, srv


def test_poisson_mix_bitexact_and_alone_equals_mix():
    from govid_amd.codec import session_bytes

    streams = _streams()
    want = {s.name: [_crc(p) for p in (lambda d: [d.decode(x) for x in s.packets])(O.OracleCodec())]
            for s in streams}
    reqs = poisson_requests(streams, 24, rate_per_s=200.0, seed=11, min_len=5, max_len=40)
    budget = 4 * session_bytes(320, 180, 4, 4)
    print(f"serving {len(reqs)} requests, budget {budget} bytes", flush=True)
    got, stats, srv = _serve(reqs, budget)
    print(stats, flush=True)
    assert stats["errors"] == 0 and srv.peak_bytes <= budget
    assert stats["max_active_sessions"] > 1
    for r in reqs:
        seqs = got[r.rid]
        assert [s for s, _, _ in seqs] == list(range(len(r.frame_order())))
        for seq, idx, crc in seqs:
            assert crc == want[r.stream.name][idx], (r.rid, seq, idx)
    # the same requests alone, one server run each
    for r in reqs[:4]:
        r.arrival = 0.0
        alone, _, _ = _serve([r], budget)
        assert alone[r.rid] == got[r.rid], r.rid
