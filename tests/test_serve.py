"""The serving loop's host logic on the CPU: admission under the HBM budget,
continuous batching, loop restarts and seeks (flush only once a session has
drained), per-request ordering, and results equal to straight-through
decoding.  The engine here is a stand-in with the batch engine's interface
whose sessions are oracle codecs."""
import zlib
from dataclasses import dataclass

import oracle as O
from govid_amd.serve import Request, Server, Stream, poisson_requests
from helpers import clip_kw


@dataclass
class _Done:
    session: int
    tag: int
    status: int
    frame: object


class OracleEngine:
    """Same calls as govid_amd.codec.Batch; each session is an OracleCodec."""

    def __init__(self, window):
        self.window = window
        self.sessions = {}
        self.queues = {}
        self.next_id = 1
        self.flushes = 0
        self.steps = 0

    def open(self, ring, window):
        sid = self.next_id
        self.next_id += 1
        self.sessions[sid] = O.OracleCodec(ring=ring)
        self.queues[sid] = []
        return sid

    def submit(self, sid, pkt, dev=None, tag=0):
        self.queues[sid].append((pkt, tag))

    def flush(self, sid):
        assert not self.queues[sid], "flush with packets still queued would drop them"
        self.flushes += 1
        self.sessions[sid].flush()

    def pending(self, sid=0):
        return sum(len(q) for q in self.queues.values())

    def step(self, max_out=4096):
        self.steps += 1
        out = []
        for sid, q in self.queues.items():
            take, self.queues[sid] = q[:self.window], q[self.window:]
            for pkt, tag in take:
                try:
                    out.append(_Done(sid, tag, 0, self.sessions[sid].decode(pkt)))
                except O.DecodeError as e:
                    out.append(_Done(sid, tag, e.status, None))
        return out

    def close_session(self, sid):
        assert not self.queues[sid]
        del self.sessions[sid], self.queues[sid]


def _streams():
    out = []
    for wl, w, h, n, kw in (("synthpan", 64, 48, 12, dict(qp=22, gop=4)),
                            ("rects", 96, 64, 10, dict(qp=22, gop=5, tile_cols=2, tile_rows=2)),
                            ("parallax", 64, 48, 9, dict(qp=28, gop=9, altref=4))):
        c = clip_kw(wl, w, h, n, threads=4, **kw)
        out.append(Stream(f"{wl}{w}x{h}", w, h, c.packets))
    return out


def _straight_crcs(stream):
    dec = O.OracleCodec()
    return [tuple(zlib.crc32(p.tobytes()) for p in dec.decode(pk)) for pk in stream.packets]


def test_mix_matches_straight_through_and_respects_budget():
    streams = _streams()
    want = {s.name: _straight_crcs(s) for s in streams}
    reqs = poisson_requests(streams, 20, rate_per_s=1e6, seed=7, min_len=3, max_len=25)
    assert any(r.seek_at is not None for r in reqs) and any(r.count > len(r.stream.packets) for r in reqs)
    got = {}

    def on_frame(fd):
        assert fd.status == 0
        got.setdefault(fd.rid, []).append((fd.seq, fd.index, tuple(zlib.crc32(p.tobytes()) for p in fd.frame)))

    cost = lambda w, h, ring, window: w * h  # noqa: E731
    budget = 3 * 96 * 64
    eng = OracleEngine(window=3)
    srv = Server(eng, budget, window=3, ring=3, session_bytes=cost, on_frame=on_frame)
    st = srv.run(reqs)
    assert st["errors"] == 0 and st["requests"] == 20
    assert srv.peak_bytes <= budget
    assert 1 < st["max_active_sessions"] <= 3 * 96 * 64 // (64 * 48)
    assert eng.flushes >= 1
    for r in reqs:
        seqs = got[r.rid]
        assert [s for s, _, _ in seqs] == list(range(len(r.frame_order())))
        for seq, idx, crc in seqs:
            assert idx == r.frame_order()[seq]
            assert crc == want[r.stream.name][idx], (r.rid, seq, idx)


def test_request_larger_than_budget_is_refused():
    s = _streams()[0]
    srv = Server(OracleEngine(2), 10, session_bytes=lambda *a: 100)
    try:
        srv.run([Request(0, s, 0, 3, 0.0)])
    except ValueError:
        return
    raise AssertionError("expected a ValueError")


def test_frame_order_loop_and_seek():
    s = Stream("x", 4, 4, [bytes([0])] + [bytes([1])] * 4)
    assert Request(0, s, 0, 7).frame_order() == [0, 1, 2, 3, 4, 0, 1]
    assert Request(0, s, 0, 5, seek_at=2, seek_key=0).frame_order() == [0, 1, 0, 1, 2]
