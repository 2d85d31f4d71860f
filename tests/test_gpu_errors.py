"""Error semantics of the GPU decoder through the C ABI, against the oracle.

Go's decoder stops a tile at the macroblock on which its rANS stream fails,
with the earlier macroblocks reconstructed (decode_intra.go:121-187); a
failing frame leaves ref and the ring image alone, and a failing golden
override leaves golden partly rewritten (codec.go:162-192).  These tests break
tiles part-way through (a tile stream cut from the front, which the decoder
reads last, or one byte inverted) in key frames, inter frames and golden
overrides of 2x2-tile clips, and check every later frame against the oracle,
plainly and through the batch engine (where a failure inside a step's window
makes the engine reconstruct the session's later frames again).

Then the host-side rejections of codec_test.go:123-175 and rans.go:396-471
(sentinel classes, the 8K pixel bound, a tile grid that does not match the
frame, table modes a key frame may not use, frequency sums), the ring reuse
of codec_test.go:58-74 with the slot a failed decode consumes
(codec.go:84-88), and seeded packet mutations in the shape of
fuzz_test.go:47-105 (cold and after a key frame).
"""
import numpy as np
import pytest

import oracle as O
from helpers import (assert_planes_equal, clip_kw, drop_front, edit_tables, flip_byte, join_packet, join_payload,
                     parse_table_blob, split_packet, split_payload)

pytestmark = pytest.mark.gpu


def _oracle_run(packets, ring=3):
    ref = O.OracleCodec(ring=ring)
    out = []
    for p in packets:
        try:
            out.append((0, ref.decode(p)))
        except O.DecodeError as e:
            out.append((e.status, None))
    ref.close()
    return out


def _plain_run(packets, ring=3, dump=False):
    """GPU statuses and planes; with dump, the tile statuses of every
    failed decode (the entropy kernel's per-tile outcome)."""
    from govid_amd.codec import Codec, EbivError

    gpu = Codec(ring=ring)
    out, mid = [], []
    for p in packets:
        try:
            out.append((0, gpu.decode(p).planes()))
        except EbivError as e:
            out.append((_status(e), None))
            if dump:
                _, _, st = gpu.debug_dump(0)
                mid.append(st)
    gpu.close()
    return out, mid


def _status(e):
    from govid_amd import codec as K

    return {K.ErrCorrupt: 1, K.ErrUnsupported: 2, K.ErrNoKeyframe: 3, K.ErrDimensions: 4}[type(e)]


def _batch_run(packets, ring, window):
    from govid_amd.codec import Batch

    b = Batch()
    sid = b.open(ring=ring, window=window)
    b.submit_many(sid, packets, tags=list(range(len(packets))))
    out = {}
    while b.pending():
        for c in b.step():
            out[c.tag] = (c.status, c.frame.planes() if c.status == 0 else None)
    b.close()
    return [out[i] for i in range(len(packets))]


def _compare(got, want, label):
    assert len(got) == len(want)
    for i, ((gs, gp), (ws, wp)) in enumerate(zip(got, want)):
        assert gs == ws, f"{label} frame {i}: status {gs}, oracle {ws}"
        if ws == 0:
            assert_planes_equal(gp, wp, f"{label} frame {i}")


def _noise_clip():
    return list(clip_kw("noisepan", 160, 128, 9, qp=28, gop=6, tile_cols=2, tile_rows=2, threads=4).packets)


def _override_clip():
    return list(clip_kw("parallax", 256, 192, 10, qp=20, gop=5, tile_cols=2, tile_rows=2, altref=4,
                        threads=4).packets)


CUTS = [("cut25", lambda p, t, pl: drop_front(p, t, 0.25, pl)),
        ("cut50", lambda p, t, pl: drop_front(p, t, 0.50, pl)),
        ("cut75", lambda p, t, pl: drop_front(p, t, 0.75, pl)),
        ("flip30", lambda p, t, pl: flip_byte(p, t, 0.30, pl)),
        ("flip70", lambda p, t, pl: flip_byte(p, t, 0.70, pl))]


@pytest.mark.parametrize("frame", [0, 2, 6])
def test_mid_tile_failures_key_and_inter(frame):
    """A tile of a key frame (0, 6) or an inter frame (2) breaks part-way;
    every frame of the clip, the failing one's status included, matches the
    oracle, plainly and batched (window 4, the failure inside the window)."""
    base = _noise_clip()
    assert base[frame][0] & 3 == (0 if frame in (0, 6) else 1)
    mid_tile = 0
    for name, cut in CUTS:
        for tile in (0, 3):
            pk = list(base)
            pk[frame] = cut(pk[frame], tile, 0)
            want = _oracle_run(pk)
            label = f"frame {frame} tile {tile} {name}"
            print(label, flush=True)
            got, sts = _plain_run(pk, dump=True)
            _compare(got, want, label + " plain")
            for st in sts:  # the failing tile stopped after some macroblocks
                bad = [s for s in st[:4] if s["err"]]
                mid_tile += sum(1 for s in bad if 0 < s["mbs_done"] < 20)
            _compare(_batch_run(pk, 4, 4), want, label + " batched")
    assert mid_tile >= 6, f"only {mid_tile} failures stopped part-way through a tile"


@pytest.mark.parametrize("frame", [0, 5])
def test_golden_override_fails_part_way(frame):
    """The override of key frame 0 or 5 breaks part-way: the key returns an
    error, ref stays the previous one and golden is partly rewritten; later
    inter frames predict from both (codec.go:176-190)."""
    base = _override_clip()
    assert len(split_packet(base[frame])[1]) == 2, "no golden override on this key frame"
    for name, cut in CUTS:
        for tile in (1, 2):
            pk = list(base)
            pk[frame] = cut(pk[frame], tile, 1)
            want = _oracle_run(pk)
            label = f"override of frame {frame} tile {tile} {name}"
            print(label, flush=True)
            got, _ = _plain_run(pk)
            _compare(got, want, label + " plain")
            _compare(_batch_run(pk, 4, 5), want, label + " batched")


def test_key_fails_part_way_before_its_override():
    """The key payload of a key frame with an override breaks part-way: Go
    never reaches the override, so its tables stay unparsed and the next key
    frame's tables start from the history the failing key left."""
    base = _override_clip()
    pk = list(base)
    pk[5] = drop_front(pk[5], 2, 0.5, 0)
    want = _oracle_run(pk)
    got, _ = _plain_run(pk)
    _compare(got, want, "key payload cut")
    _compare(_batch_run(pk, 4, 6), want, "key payload cut batched")


# ---- host-side rejections --------------------------------------------------------

def _decode_status(packets, ring=3):
    """Status of each packet decoded in order by the GPU codec and the oracle."""
    got, _ = _plain_run(packets, ring)
    want = _oracle_run(packets, ring)
    _compare(got, want, "sequence")
    return [s for s, _ in got]


def _raw_key(w, h):
    n = w * h + 2 * ((w + 1) // 2) * ((h + 1) // 2)
    return bytes([0x00, w & 255, w >> 8, h & 255, h >> 8, 0]) + bytes(range(256)) * (n // 256) + bytes(n % 256)


def test_sentinels_of_bad_packets():
    """TestDecodeRejectsBadPackets and the raw-length / resolution-change
    cases (codec_test.go:123-175), and the 8192x4320 pixel bound."""
    assert _decode_status([b""]) == [1]
    assert _decode_status([bytes([1])]) == [3]
    assert _decode_status([bytes([0, 1, 0])]) == [1]
    raw = _raw_key(16, 16)
    assert _decode_status([raw[:6]]) == [1]
    assert _decode_status([raw + b"\0"]) == [1]
    assert _decode_status([raw, _raw_key(32, 32), raw]) == [0, 4, 0]
    for w, h in ((8193, 4320), (65535, 65535), (8192, 4321)):
        hdr = bytes([0x04, w & 255, w >> 8, h & 255, h >> 8, 0])
        assert _decode_status([hdr + bytes(64)]) == [2], (w, h)
    assert _decode_status([bytes([0x04, 0, 0, 16, 0, 0]) + bytes(16)]) == [1]  # zero width
    assert _decode_status([bytes([0x04, 16, 0, 16, 0, 0x10]) + bytes(16)]) == [2]  # chroma format
    assert _decode_status([bytes([0x04, 16, 0, 16, 0, 0x01]) + bytes(16)]) == [2]  # bit depth
    assert _decode_status([bytes([0x1D])]) == [3]  # inter frame (unknown coding) before any key
    coded = list(clip_kw("synthpan", 64, 48, 3, qp=22, gop=3).packets)
    assert _decode_status([coded[0], bytes([0x1D]) + coded[1][1:], coded[1]]) == [0, 2, 0]
    assert _decode_status([coded[0], bytes([0x21]) + coded[1][1:], coded[1]]) == [0, 1, 0]  # override on inter
    # a coded key frame of another size mid-stream
    other = clip_kw("synthpan", 80, 48, 1, qp=22, gop=1).packets[0]
    assert _decode_status(coded[:2] + [other, coded[2]]) == [0, 0, 4, 0]


def test_tile_grid_that_does_not_match_the_frame():
    """newTileGrid clamps the grid to the macroblock grid; a header naming more
    tile columns than the frame has macroblock columns then disagrees with
    the tile count (decode_intra.go:60-63) -> ErrCorrupt."""
    pk = list(clip_kw("synthpan", 64, 48, 3, qp=22, gop=3).packets)  # 4x3 macroblocks, one tile
    head, pls = split_packet(pk[0])
    qp, cols, rows, tables, tiles, tr = split_payload(pls[0])
    bad = join_packet(head, [join_payload(qp, 5, 1, tables, tiles + [tiles[0]] * 4, tr)])
    assert _decode_status([bad, pk[0], pk[1]]) == [1, 0, 0]
    short = join_packet(head, [join_payload(qp, 2, 2, tables, tiles, tr)])  # too few tile streams
    assert _decode_status([short, pk[0]]) == [1, 0]


def test_table_rejects_and_wrapped_frequencies():
    """parseTablesInto (rans.go:396-471): a key frame may not reference or
    delta a previous table; a context whose frequencies do not sum to 4096 is
    corrupt; a full-table value read as uint32(uvarint) wraps, so 2^32 + f
    decodes as f; values above 4096 whose uint32 sum wraps to 4096 are
    rejected (Go would index out of range there)."""
    from govid_amd.codec import Codec, ErrCorrupt

    pk = list(clip_kw("rects", 96, 64, 4, qp=22, gop=4).packets)
    full = [c for c, (m, _) in enumerate(parse_table_blob(split_payload(split_packet(pk[0])[1][0])[3])) if m == 2]
    c0 = full[0]

    def set_mode(mode, vals=None):
        def fn(ent):
            ent = list(ent)
            ent[c0] = (mode, ent[c0][1] if vals is None else vals(ent[c0][1]))
            return ent
        return fn

    cases = {
        "same on key": edit_tables(pk[0], set_mode(1, lambda v: [])),
        "delta on key": edit_tables(pk[0], set_mode(3, lambda v: [0] * len(v))),
        "sum 4095": edit_tables(pk[0], set_mode(2, lambda v: [v[0] - 1] + v[1:] if v[0] > 1 else [v[0] + 1] + v[1:])),
        "wrapped above 4096": edit_tables(pk[0], set_mode(2, lambda v: [(1 << 32) - 1, 4097] + [0] * (len(v) - 2))),
        "unknown mode": edit_tables(pk[0], set_mode(7, lambda v: [])),
    }
    for name, bad in cases.items():
        print(name, flush=True)
        assert _decode_status([bad, pk[0], pk[1]]) == [1, 0, 0], name
    gpu = Codec()
    with pytest.raises(ErrCorrupt, match="previous table"):
        gpu.decode(cases["same on key"])
    with pytest.raises(ErrCorrupt, match="sum to 4096"):
        gpu.decode(cases["sum 4095"])
    gpu.close()
    # 2^32 + f reads as f: the frame decodes exactly as the original
    wrapped = edit_tables(pk[0], set_mode(2, lambda v: [v[0] + (1 << 32)] + v[1:]))
    assert len(wrapped) > len(pk[0])
    got, _ = _plain_run([wrapped] + pk[1:])
    want = _oracle_run(pk)
    _compare(got, want, "wrapped full table")


def test_ring_reuse_and_the_slot_a_failure_consumes():
    """TestCodecReusesRingBuffers (codec_test.go:58-74): after ring decodes
    the same host plane buffers come back; a failed decode still consumes
    its ring slot (codec.go:84-88), so the next frame lands one slot on."""
    import ctypes as C

    from govid_amd import capi
    from govid_amd.codec import Codec
    from helpers import zero_tile

    ring = 3
    pk = list(clip_kw("synthpan", 64, 48, 8, qp=22, gop=8).packets)
    gpu = Codec(ring=ring)
    L = gpu._L
    slots, ptrs = [], []
    seq = pk[:4] + [zero_tile(pk[4])] + pk[5:]
    for p in seq:
        fr = capi.Frame()
        rc = L.ebiv_codec_decode(gpu._h, p, len(p), 0, C.byref(fr))
        slots.append(fr.ring_slot if rc == 0 else None)
        ptrs.append(fr.y if rc == 0 else None)
    gpu.close()
    assert slots == [0, 1, 2, 0, None, 2, 0, 1]
    assert ptrs[3] == ptrs[0] and ptrs[5] == ptrs[2] and ptrs[6] == ptrs[0] and ptrs[7] == ptrs[1]
    assert len({ptrs[0], ptrs[1], ptrs[2]}) == 3
    _decode_status(seq)


def _mutate(rng, p: bytes) -> bytes:
    b = bytearray(p)
    kind = int(rng.integers(5))
    if kind == 0 and b:  # bit flips
        for _ in range(int(rng.integers(1, 9))):
            i = int(rng.integers(len(b)))
            b[i] ^= 1 << int(rng.integers(8))
    elif kind == 1 and b:  # truncation
        b = b[:int(rng.integers(len(b)))]
    elif kind == 2:  # inserted bytes
        i = int(rng.integers(len(b) + 1))
        b[i:i] = rng.integers(0, 256, int(rng.integers(1, 9))).astype(np.uint8).tobytes()
    elif kind == 3 and len(b) > 1:  # deleted range
        i = int(rng.integers(len(b)))
        del b[i:i + int(rng.integers(1, 17))]
    else:  # random bytes over a range
        i = int(rng.integers(len(b) + 1))
        n = int(rng.integers(1, 33))
        b[i:i + n] = rng.integers(0, 256, n).astype(np.uint8).tobytes()
    return bytes(b)


def test_seeded_mutations_cold_and_warm():
    """FuzzDecodePacket's two uses (fuzz_test.go:47-87) over seeded mutations
    of a small stream with a golden override, inter, skip and golden paths:
    each mutated packet decoded by a fresh codec, and after the stream's key
    frame; status and planes match the oracle.  (The host parse of the same
    mutations also runs under AddressSanitizer on the CPU.)"""
    pk = list(clip_kw("synthframe", 64, 48, 6, qp=20, gop=4, altref=4).packets)
    rng = np.random.default_rng(47)
    n = 0
    for case in range(120):
        src = pk[int(rng.integers(len(pk)))]
        m = _mutate(rng, src)
        if case % 20 == 0:
            print(f"mutation case {case}", flush=True)
        _decode_status([m])
        _decode_status([pk[0], m, pk[1]])
        n += 1
    assert n == 120
