"""The benchmark workloads beyond the 1080p clip, decoded on the GPU and
compared frame by frame with the oracle through the C ABI:

- the mixed batch: 640x360, 1280x720 and 1920x1080 streams with 1x1, 2x2 and
  4x3 tile grids, q 18/22/28, decoded together by one batch engine;
- the entropy-bound stream: 1920x1080 uniform noise (sigma 20) under random
  global motion, q12, gop 15, 4x3 tiles, decoded plain and as 8 batched
  sessions from staggered key frames;
- the serving mix: Poisson plays, seeks and loop restarts over 640x360,
  1280x720 and 1920x1080 single-tile streams and a 3840x2160 stream with 8x6
  tiles, under an HBM budget that makes requests wait; each request's frames
  are also decoded alone and must equal the mix's.

Streams come from bench_data/make_streams.py (the generated files when present,
else encoded here from the same seeded specs), cut to their first frames so
that the oracle finishes in seconds.
"""
import pytest

from helpers import bench_stream, crc_planes, mixed_names, oracle_crcs, stream_spec

pytestmark = pytest.mark.gpu


def _mixed_pick():
    """One stream per (size, tile grid) pair of the mixed set, in spec order."""
    seen, out = set(), []
    for nm in mixed_names():
        sp = stream_spec(nm)
        key = (sp["w"], sp["enc"]["tile_cols"])
        if key not in seen:
            seen.add(key)
            out.append(nm)
    return out


def _check_batch(streams, window, ring=None):
    """streams: [(label, packets, start)]: one session each from frame `start`.
    Every completion is compared with the oracle's decode of that frame."""
    from govid_amd.codec import Batch

    b = Batch()
    sids = []
    for k, (label, pk, start) in enumerate(streams):
        sid = b.open(ring=ring or window, window=window)
        sids.append(sid)
        order = list(range(start, len(pk))) + list(range(0, start))
        b.submit_many(sid, [pk[i] for i in order], tags=[(k << 32) | i for i in order])
    wants = [oracle_crcs(tuple(pk)) for _, pk, _ in streams]
    n = steps = 0
    while b.pending():
        done = b.step()
        steps += 1
        assert done
        for c in done:
            k, i = c.tag >> 32, c.tag & 0xFFFFFFFF
            assert c.session == sids[k]
            st, crc = wants[k][i]
            assert c.status == st, f"{streams[k][0]} frame {i}: status {c.status}, oracle {st}"
            if st == 0:
                assert crc_planes(c.frame.planes()) == crc, f"{streams[k][0]} frame {i}: planes differ from the oracle"
            n += 1
        print(f"  step {steps}: {n} frames checked", flush=True)
    b.close()
    return n


def test_mixed_batch_sizes_and_grids():
    """Config 3's mix: every size with every tile grid the set draws, one
    session each, window 8 (the first 24 frames of each stream)."""
    names = _mixed_pick()
    streams = []
    for nm in names:
        pk, w, h = bench_stream(nm, 24)
        sp = stream_spec(nm)
        label = f"{nm} {w}x{h} {sp['enc']['tile_cols']}x{sp['enc']['tile_rows']} q{sp['enc']['qp']}"
        print(label, flush=True)
        streams.append((label, pk, 0))
    assert {stream_spec(n)["w"] for n in names} == {640, 1280, 1920}
    assert {stream_spec(n)["enc"]["tile_cols"] for n in names} == {1, 2, 4}
    assert _check_batch(streams, window=8) == sum(len(s[1]) for s in streams)


def _noise():
    pk, w, h = bench_stream("noisepan_1080p", 30)
    keys = [i for i, p in enumerate(pk) if p[0] & 3 == 0]
    assert (w, h) == (1920, 1080) and len(pk) >= 16 and len(keys) >= 2, (len(pk), keys)
    return pk, keys


def test_entropy_bound_stream_plain():
    """Config 4's stream through the plain drop-in, every frame."""
    from govid_amd.codec import Codec

    pk, _ = _noise()
    want = oracle_crcs(tuple(pk))
    gpu = Codec(ring=3)
    for i, p in enumerate(pk):
        f = gpu.decode(p)
        assert crc_planes(f.planes()) == want[i][1], f"noise frame {i}"
    print(f"{len(pk)} noise frames bit-exact", flush=True)
    gpu.close()


def test_entropy_bound_stream_batched():
    """Config 4's stream as 8 sessions from staggered key frames, window 8."""
    pk, keys = _noise()
    streams = [(f"noise session {s}", pk, keys[s % len(keys)]) for s in range(8)]
    assert _check_batch(streams, window=8) == 8 * len(pk)


def _serving_streams(frames):
    from govid_amd.serve import Stream

    out = []
    for nm in ("serve_360p", "serve_720p", "serve_1080p", "serve_2160p"):
        pk, w, h = bench_stream(nm, frames)
        out.append(Stream(nm, w, h, list(pk)))
    return out


def test_serving_mix_with_2160p_tiles():
    """Config 5's sizes and encoding (single-tile q22 gop 30 at 360p-1080p,
    8x6 tiles at 2160p) under Poisson arrivals with seeks and loop restarts;
    the budget admits about two 2160p sessions at a time."""
    from govid_amd.codec import Batch, session_bytes
    from govid_amd.serve import Server, poisson_requests

    streams = _serving_streams(40)
    assert stream_spec("serve_2160p")["enc"]["tile_cols"] == 8 and streams[3].width == 3840
    want = {s.name: [c for _, c in oracle_crcs(tuple(s.packets))] for s in streams}
    reqs = poisson_requests(streams, 14, rate_per_s=200.0, seed=15, min_len=10, max_len=60)
    assert any(r.seek_at is not None for r in reqs) and any(len(r.frame_order()) > 40 for r in reqs)
    window = 2
    budget = 2 * session_bytes(3840, 2160, window, window) + session_bytes(1920, 1080, window, window)

    def serve(rs):
        got = {}

        def on_frame(fd):
            assert fd.status == 0, (fd.rid, fd.seq)
            got.setdefault(fd.rid, []).append((fd.seq, fd.index, crc_planes(fd.frame.planes())))

        b = Batch()
        srv = Server(b, budget, window=window, ring=window, on_frame=on_frame)
        stats = srv.run(rs, max_seconds=300)
        b.close()
        return got, stats, srv

    print(f"serving {len(reqs)} requests, budget {budget} bytes", flush=True)
    got, stats, srv = serve(reqs)
    print(stats, flush=True)
    assert stats["errors"] == 0 and srv.peak_bytes <= budget
    assert stats["max_active_sessions"] > 1 and stats["admission_wait_p99_ms"] > 0
    for r in reqs:
        seqs = got[r.rid]
        assert [s for s, _, _ in seqs] == list(range(len(r.frame_order())))
        for seq, idx, crc in seqs:
            assert crc == want[r.stream.name][idx], (r.rid, r.stream.name, seq, idx)
    # alone == mix: a seek, a loop restart and a 2160p play, each served alone
    picks = [next(r for r in reqs if r.seek_at is not None),
             next(r for r in reqs if len(r.frame_order()) > len(r.stream.packets)),
             next(r for r in reqs if r.stream.width == 3840)]
    for r in picks:
        r.arrival = 0.0
        alone, _, _ = serve([r])
        assert alone[r.rid] == got[r.rid], r.rid
