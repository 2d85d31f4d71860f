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


def _stream_specs():
    import os
    import sys

    bd = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench_data")
    if bd not in sys.path:
        sys.path.insert(0, bd)
    import make_streams

    return make_streams


@functools.lru_cache(maxsize=32)
def bench_stream(name: str, frames: int):
    """The first `frames` packets of benchmark stream `name`
    (bench_data/make_streams.py): read from the generated file when it
    exists, else encoded here from the same seeded spec, cut to `frames`
    frames.  Returns (packets, width, height)."""
    import os

    from govid_amd.container import Reader

    ms = _stream_specs()
    spec = ms.STREAMS[name]
    path = ms.path_of(name)
    if os.path.exists(path):
        return tuple(Reader.open(path).packets()[:frames]), spec["w"], spec["h"]
    n = min(frames, spec["frames"])
    c = clip(spec["workload"], spec["w"], spec["h"], n, spec["seed"], threads=8, **spec["enc"])
    return tuple(c.packets), spec["w"], spec["h"]


def stream_spec(name: str) -> dict:
    return _stream_specs().STREAMS[name]


def mixed_names() -> list:
    return sorted(k for k in _stream_specs().STREAMS if k.startswith("mixed_"))


def crc_planes(planes):
    import zlib

    return tuple(zlib.crc32(np.ascontiguousarray(p).tobytes()) for p in planes)


@functools.lru_cache(maxsize=32)
def oracle_crcs(packets: tuple, threads: int = 8):
    """Per-packet (status, plane CRCs) of the oracle's straight-through
    decode of `packets`."""
    import oracle as O

    dec = O.OracleCodec(ring=3, threads=threads)
    out = []
    for p in packets:
        try:
            out.append((0, crc_planes(dec.decode(p))))
        except O.DecodeError as e:
            out.append((e.status, None))
    dec.close()
    return tuple(out)


# ---- packet surgery (frame header format.go:376-409, override split
# codec.go:162-173, payload frame.go:197-274, tables rans.go:396-471) ----------

ALPHABETS = [3, 10, 10, 3, 2, 4, 8, 2, 16, 16, 16, 2] + [7] * 60  # model.go:195-213


def put_uvarint(v: int) -> bytes:
    out = bytearray()
    while v >= 0x80:
        out.append((v & 0x7F) | 0x80)
        v >>= 7
    out.append(v)
    return bytes(out)


def split_packet(pkt: bytes):
    """(header bytes, [payload bytes...]): one payload, or the key's and the
    override's for a key frame with a golden override."""
    typ, gov = pkt[0] & 3, (pkt[0] >> 5) & 1
    hlen = 6 if typ == 0 else 1
    if not gov:
        return pkt[:hlen], [pkt[hlen:]]
    kl, p2 = _uvarint(pkt, hlen)
    return pkt[:hlen], [pkt[p2:p2 + kl], pkt[p2 + kl:]]


def join_packet(head: bytes, payloads) -> bytes:
    if len(payloads) == 1:
        return head + payloads[0]
    return head + put_uvarint(len(payloads[0])) + payloads[0] + payloads[1]


def split_payload(body: bytes):
    """(qp, cols, rows, table blob, [tile bytes], trailing bytes)."""
    qp = body[0]
    cols, pos = _uvarint(body, 1)
    rows, pos = _uvarint(body, pos)
    tlen, pos = _uvarint(body, pos)
    tables = body[pos:pos + tlen]
    pos += tlen
    tiles = []
    for _ in range(cols * rows):
        ln, pos = _uvarint(body, pos)
        tiles.append(body[pos:pos + ln])
        pos += ln
    return qp, cols, rows, tables, tiles, body[pos:]


def join_payload(qp, cols, rows, tables, tiles, trailing=b"") -> bytes:
    out = bytearray([qp]) + put_uvarint(cols) + put_uvarint(rows) + put_uvarint(len(tables)) + tables
    for t in tiles:
        out += put_uvarint(len(t)) + t
    return bytes(out + trailing)


def edit_tile(pkt: bytes, tile: int, fn, payload: int = 0) -> bytes:
    """The packet with tile `tile` of payload `payload` replaced by fn(bytes)
    (its length field rewritten to match)."""
    head, pls = split_packet(pkt)
    qp, cols, rows, tables, tiles, tr = split_payload(pls[payload])
    tiles[tile] = fn(tiles[tile])
    pls[payload] = join_payload(qp, cols, rows, tables, tiles, tr)
    return join_packet(head, pls)


def drop_front(pkt: bytes, tile: int, frac: float, payload: int = 0) -> bytes:
    """Removes the first `frac` of a tile stream.  rANS reads a tile from its
    end towards its start (rans.go:237-263), so the decode runs normally
    until it needs the removed bytes and then underruns part-way through."""
    return edit_tile(pkt, tile, lambda t: t[int(len(t) * frac):], payload)


def flip_byte(pkt: bytes, tile: int, frac: float, payload: int = 0) -> bytes:
    """Inverts the byte the decode reaches after about `frac` of the tile."""
    def fn(t):
        b = bytearray(t)
        p = max(0, min(len(b) - 1, len(b) - 1 - int(len(b) * frac)))
        b[p] ^= 0xFF
        return bytes(b)
    return edit_tile(pkt, tile, fn, payload)


def parse_table_blob(blob: bytes):
    """[(mode, [values])] for the 72 contexts of a table block: full tables
    as uvarints, deltas as zig-zag varints (kept raw, unsigned)."""
    out, pos = [], 0
    for c in range(72):
        mode = blob[pos]
        pos += 1
        vals = []
        if mode in (2, 3):
            for _ in range(ALPHABETS[c]):
                v, pos = _uvarint(blob, pos)
                vals.append(v)
        out.append((mode, vals))
    assert pos == len(blob)
    return out


def build_table_blob(entries) -> bytes:
    out = bytearray()
    for mode, vals in entries:
        out.append(mode)
        for v in vals:
            out += put_uvarint(v)
    return bytes(out)


def edit_tables(pkt: bytes, fn, payload: int = 0) -> bytes:
    """The packet with its table block replaced by build(fn(parsed entries))."""
    head, pls = split_packet(pkt)
    qp, cols, rows, tables, tiles, tr = split_payload(pls[payload])
    pls[payload] = join_payload(qp, cols, rows, build_table_blob(fn(parse_table_blob(tables))), tiles, tr)
    return join_packet(head, pls)
