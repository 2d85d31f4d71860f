"""Batched decode (many sessions, several frames per session per step) vs the
oracle, completion by completion, through the C ABI.

Every session must behave exactly as its own ebiv.Codec would
(codec.go:70-207): the same frames, the same errors, the same state after an
error — including a failing frame inside a step's window, whose later frames
decode against the reference the failure left in place.
"""
import numpy as np
import pytest

import oracle as O
from helpers import assert_planes_equal, clip_kw, zero_tile

pytestmark = pytest.mark.gpu

ERR = {1: "ErrCorrupt", 2: "ErrUnsupported", 3: "ErrNoKeyframe", 4: "ErrDimensions"}


def _oracle_outcomes(packets, ring, flush_at=()):
    ref = O.OracleCodec(ring=ring)
    out = []
    for i, p in enumerate(packets):
        if i in flush_at:
            ref.flush()
        try:
            out.append((0, ref.decode(p)))
        except O.DecodeError as e:
            out.append((e.status, None))
    return out


def _run_batch(streams, max_out=4096, check_rgba=False, schedule=None, bulk=False):
    """streams: list of (packets, ring, window).  Submits everything, steps
    until drained, and checks each completion against the oracle."""
    from govid_amd.codec import Batch

    b = Batch()
    if schedule:
        b.set_schedule(**schedule)
        assert all(b.schedule()[k] == v for k, v in schedule.items())
    want, sids = [], []
    for k, (packets, ring, window) in enumerate(streams):
        sid = b.open(ring=ring, window=window)
        sids.append(sid)
        want.append(_oracle_outcomes(packets, max(ring, window)))
        if bulk:  # one call per stream (ebiv_batch_submit_many)
            b.submit_many(sid, packets, tags=[(k << 32) | i for i in range(len(packets))])
        else:
            for i, p in enumerate(packets):
                b.submit(sid, p, tag=(k << 32) | i)
    seen = [0] * len(streams)
    steps = 0
    while b.pending():
        done = b.step(max_out)
        steps += 1
        assert done, "a step with pending packets completed nothing"
        for c in done:
            k, i = c.tag >> 32, c.tag & 0xFFFFFFFF
            assert c.session == sids[k]
            assert i == seen[k], f"session {k}: completion {i} out of order (expected {seen[k]})"
            seen[k] += 1
            st, planes = want[k][i]
            print(f"  step {steps} session {k} frame {i}: status {c.status} (oracle {st})", flush=True)
            assert c.status == st, f"session {k} frame {i}: status {c.status} != oracle {ERR.get(st, st)}"
            if st == 0:
                assert_planes_equal(c.frame.planes(), planes, f"session {k} frame {i}")
                if check_rgba:
                    assert np.array_equal(c.frame.rgba(), O.rgba(*planes)), f"session {k} frame {i} rgba"
    assert seen == [len(s[0]) for s in streams]
    b.close()
    return steps


def _clip(wl, w, h, n, **kw):
    return clip_kw(wl, w, h, n, threads=4, **kw).packets


def test_sessions_of_mixed_geometry_and_windows():
    streams = [
        (_clip("synthpan", 64, 48, 9, qp=22, gop=4), 3, 4),
        (_clip("rects", 96, 64, 10, qp=22, gop=10, tile_cols=4, tile_rows=3), 3, 3),
        (_clip("synthpan", 37, 21, 5, qp=22, gop=5), 3, 1),
        (_clip("parallax", 64, 48, 9, qp=28, gop=9, altref=4), 4, 5),
        (_clip("synthframe", 64, 48, 3, qp=20, gop=1), 2, 2),
    ]
    print("mixed sessions", flush=True)
    _run_batch(streams, check_rgba=True)


@pytest.mark.parametrize("schedule", [
    dict(intra_waves_key=1, intra_waves_few=1, intra_waves_many=1),
    dict(intra_waves_key=5, intra_waves_few=3, intra_waves_many=7, intra_waves_key_many=2, many_tiles=1),
    dict(intra_waves_key=16, intra_waves_few=16, intra_waves_many=16, many_tiles=1000),
    dict(recon_lanes=2),
    dict(recon_lanes=4, intra_waves_few=3),
    dict(pipeline=2),
    dict(pipeline=2, recon_lanes=2),
    dict(intra_cells=1),
    dict(intra_cells=1, intra_waves_key=16, intra_waves_key_many=16, intra_waves_few=16, intra_waves_many=16),
    dict(intra_cells=1, intra_waves_key=2, intra_waves_key_many=2, intra_waves_few=2, intra_waves_many=2,
         many_tiles=1),
    dict(key_band_rows=1),
    dict(key_band_rows=2, intra_waves_key=1, intra_waves_key_many=1),
    dict(key_band_rows=2, recon_lanes=2, many_tiles=1000),
    dict(key_band_rows=3, many_tiles=1),  # (a step of many key-frame tiles keeps them whole)
    dict(key_band_rows=128),
])
def test_launch_schedules_do_not_change_results(schedule):
    """Intra reconstruction with 1 to 16 wavefronts per tile (key frames from
    the LDS frontier, inter frames from the frame), by block levels or by
    macroblock wavefronts, key-frame tiles in bands of 1 or 2 rows or whole,
    reconstruction split over 2 or 4 streams, and
    steps run one at a time instead of the next step's entropy launch beside
    this step's reconstruction, bit for bit."""
    streams = [
        (_clip("rects", 96, 64, 8, qp=22, gop=4, tile_cols=2, tile_rows=2), 3, 4),
        (_clip("parallax", 64, 48, 7, qp=28, gop=7, altref=3), 4, 4),
        (_clip("synthframe", 80, 48, 4, qp=18, gop=2), 3, 4),
    ]
    print(f"schedule {schedule}", flush=True)
    _run_batch(streams, schedule=schedule)


@pytest.mark.parametrize("schedule", [
    dict(many_tiles=1),
    dict(many_tiles=1, intra_waves_key_many=1, intra_waves_many=2),
    dict(many_tiles=1, recon_lanes=2, intra_waves_key_many=16),
    dict(many_tiles=64, intra_waves_key=3),
    dict(many_tiles=1000),
])
def test_inter_frames_coded_intra(schedule):
    """The parallax clip's flash at frame 14 is an inter frame whose tiles are
    all intra.  In a step of at least many_tiles tiles such tiles get a launch
    of their own (LDS frontier, shaped as a key-frame launch) beside the launch
    of the round's other inter-frame tiles; with fewer tiles they run banded.
    Sessions start at different frames, so the flash falls in different
    rounds of a step, bit for bit against the oracle."""
    pk = _clip("parallax", 128, 96, 28, qp=22, gop=60, tile_cols=2, tile_rows=2)
    streams = [(pk, 3, 16), (pk[12:], 3, 8), (pk[:17], 4, 16)]
    print(f"flash clip, schedule {schedule}", flush=True)
    _run_batch(streams, schedule=schedule)


CELLS_16 = dict(intra_cells=1, intra_waves_key=16, intra_waves_key_many=16, intra_waves_few=16, intra_waves_many=16)


@pytest.mark.parametrize("case", [
    dict(w=1024, h=256, tiles=(1, 1)),   # 64 x 16 macroblocks: one row per wavefront, the widest tile
    dict(w=992, h=1024, tiles=(2, 2)),   # 31 x 32 tiles: two rows per wavefront, 992 macroblocks
    dict(w=1024, h=1024, tiles=(2, 2)),  # 32 x 32 tiles: by macroblock wavefronts (rows would overlap)
    dict(w=1040, h=64, tiles=(1, 1)),    # 65 macroblocks wide: by macroblock wavefronts
    dict(w=528, h=512, tiles=(1, 1)),    # 33 x 32 = 1056 macroblocks: by macroblock wavefronts
])
def test_intra_block_levels_at_their_tile_limits(case):
    """The block-level intra reconstruction (schedule intra_cells=1, 16
    wavefronts) on key frames and inter frames in tiles at and just past the
    sizes it takes, against the oracle."""
    w, h, (tc, tr) = case["w"], case["h"], case["tiles"]
    print(f"case {w}x{h} tiles {tc}x{tr}", flush=True)
    pk = clip_kw("parallax", w, h, 4, qp=22, gop=2, tile_cols=tc, tile_rows=tr, threads=8).packets
    _run_batch([(pk, 3, 1)], schedule=CELLS_16)


def test_intra_block_levels_on_the_bench_clip():
    """The 1080p clip's key and flash frames (30 x 23-macroblock tiles, two
    rows per wavefront) by block levels, every frame against the oracle."""
    import os

    from govid_amd.container import Reader

    path = os.path.join(os.path.dirname(os.path.dirname(__file__)), "bench_data", "parallax_1080p.ebiv")
    pk = Reader.open(path).packets()[:40]
    _run_batch([(pk, 3, 4)], schedule=CELLS_16)


def test_tuner_checks_every_schedule_and_picks_one():
    from govid_amd import tune

    pk = clip_kw("rects", 96, 64, 6, qp=22, gop=3, tile_cols=2, tile_rows=2).packets
    res = tune.tune(pk, sessions=3, window=2, passes=1)
    assert len(res["tried"]) >= 2
    assert res["best"] in [t["schedule"] for t in res["tried"]]
    assert all(t["frames"] == 3 * len(pk) for t in res["tried"])


def test_tuner_window_search_keeps_results():
    """Every candidate window decodes the same frames (the tuner compares the
    CRCs of every frame) and the chosen one is among them."""
    from govid_amd import tune

    pk = clip_kw("rects", 96, 64, 7, qp=22, gop=3, tile_cols=2, tile_rows=2).packets
    print("tuner windows 1, 2, 5", flush=True)
    res = tune.tune(pk, sessions=3, passes=1, windows=(1, 2, 5))
    assert [w["window"] for w in res["windows_tried"]] == [1, 2, 5]
    assert res["window"] in (1, 2, 5)
    assert all(w["frames"] == 3 * len(pk) for w in res["windows_tried"])


def test_bulk_submission_matches_oracle():
    streams = [
        (_clip("rects", 96, 64, 10, qp=22, gop=5, tile_cols=2, tile_rows=2), 3, 3),
        (_clip("parallax", 64, 48, 9, qp=28, gop=9, altref=4), 4, 2),
    ]
    print("bulk submit", flush=True)
    _run_batch(streams, bulk=True)


def test_window_larger_than_stream_and_max_out_limit():
    pk = _clip("synthpan", 64, 48, 6, qp=22, gop=6)
    print("window 8 over 6 frames", flush=True)
    _run_batch([(pk, 3, 8)])
    print("max_out 2", flush=True)
    assert _run_batch([(pk, 3, 4), (pk, 3, 4)], max_out=2) >= 6


@pytest.mark.parametrize("bad", [[2], [0], [2, 3], [4, 7]])
def test_failures_inside_the_window(bad):
    """A tile that underruns on the device: the frame fails, later frames in
    the same step predict from the reference Go kept (codec.go:174-190)."""
    pk = list(_clip("synthpan", 64, 48, 9, qp=22, gop=4))
    for i in bad:
        pk[i] = zero_tile(pk[i])
    print(f"corrupt frames {bad}", flush=True)
    _run_batch([(pk, 3, 5), (list(_clip("synthpan", 64, 48, 9, qp=22, gop=4)), 3, 5)])


@pytest.mark.parametrize("payload", [0, 1])
def test_golden_override_failure(payload):
    """Key frame with an alt-ref override: a failing key never reaches the
    override (its tables stay unparsed); a failing override leaves golden
    partially rewritten (codec.go:176-181)."""
    pk = list(_clip("parallax", 64, 48, 9, qp=28, gop=9, altref=4))
    from helpers import payload_spans

    assert len(payload_spans(pk[0])) == 2, "clip has no golden override"
    pk[0] = zero_tile(pk[0], 0, payload)
    print(f"corrupt payload {payload} of the key frame", flush=True)
    _run_batch([(pk, 3, 4)])


def test_host_side_errors_in_window():
    """Errors found by the host (no key frame yet, truncated header, unknown
    coding) keep their place in the completion order."""
    pk = list(_clip("synthpan", 64, 48, 6, qp=22, gop=3))
    stream = [pk[1], pk[0], pk[1][:1], pk[2], bytes([0x1C]) + pk[3][1:], pk[3], pk[4], pk[5]]
    _run_batch([(stream, 3, 4)])


def test_flush_and_seek_to_key():
    """Seek: flush the session, then continue from a later key frame."""
    from govid_amd.codec import Batch

    pk = list(_clip("synthpan", 64, 48, 8, qp=22, gop=4))
    b = Batch()
    sid = b.open(ring=3, window=4)
    ref = O.OracleCodec(ring=3)
    for p in pk[:3]:
        b.submit(sid, p)
    got = [c.frame.planes() for c in b.step()]
    for g, p in zip(got, pk[:3]):
        assert_planes_equal(g, ref.decode(p))
    b.submit(sid, pk[3])
    b.flush(sid)
    ref.flush()
    assert b.pending(sid) == 0
    for p in pk[4:]:
        b.submit(sid, p)
    done = b.step()
    assert len(done) == 4
    for c, p in zip(done, pk[4:]):
        assert c.status == 0
        assert_planes_equal(c.frame.planes(), ref.decode(p))
    b.close()


def test_bench_clip_two_sessions_window_8():
    """The benchmark's 1080p clip through the batched engine, two sessions."""
    import os

    from govid_amd.container import Reader

    path = os.path.join(os.path.dirname(os.path.dirname(__file__)), "bench_data", "parallax_1080p.ebiv")
    pk = Reader.open(path).packets()[:40]
    from govid_amd.codec import Batch

    b = Batch()
    s1, s2 = b.open(ring=8, window=8), b.open(ring=8, window=8)
    for p in pk:
        b.submit(s1, p)
    for p in pk[:24]:
        b.submit(s2, p)
    ref1, ref2 = O.OracleCodec(ring=8, threads=8), O.OracleCodec(ring=8, threads=8)
    n1 = n2 = 0
    while b.pending():
        for c in b.step():
            assert c.status == 0, c.error
            if c.session == s1:
                assert_planes_equal(c.frame.planes(), ref1.decode(pk[n1]), f"s1 frame {n1}")
                n1 += 1
            else:
                assert_planes_equal(c.frame.planes(), ref2.decode(pk[n2]), f"s2 frame {n2}")
                n2 += 1
        print(f"  {n1} + {n2} frames checked", flush=True)
    assert (n1, n2) == (40, 24)
    b.close()


def test_closing_a_session_with_a_step_in_flight():
    """A step() call also starts the next step's entropy decoding; a session
    closed before the next call gets no completions from it, and the other
    sessions' frames are unaffected."""
    from govid_amd.codec import Batch

    pa = list(_clip("synthpan", 64, 48, 9, qp=22, gop=4))
    pb = list(_clip("rects", 96, 64, 10, qp=22, gop=5, tile_cols=2, tile_rows=2))
    b = Batch()
    sa, sb = b.open(ring=3, window=3), b.open(ring=3, window=3)
    b.submit_many(sa, pa, tags=list(range(len(pa))))
    b.submit_many(sb, pb, tags=list(range(len(pb))))
    first = b.step()
    assert {c.session for c in first} == {sa, sb}
    assert b.pending(sa) == len(pa) - 3  # three in flight, three queued
    b.close_session(sa)
    ref = _oracle_outcomes(pb, 3)
    seen = []

    def check(done):  # at once: a frame stays valid for `ring` more frames
        for c in done:
            assert c.session == sb
            st, planes = ref[c.tag]
            assert c.status == st
            if st == 0:
                assert_planes_equal(c.frame.planes(), planes, f"frame {c.tag}")
            seen.append(c.tag)

    check([c for c in first if c.session == sb])
    while b.pending():
        check(b.step())
    assert seen == list(range(len(pb)))
    b.close()
