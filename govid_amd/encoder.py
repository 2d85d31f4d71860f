"""EBIV stream encoder (host C++, libebivenc) — Python wrapper.

Mirrors ebiv.NewEncoder(w, Config, opts...) (ebiv/encoder.go:128-176):
WithIntra(qp), WithGOP(gop), WithTiles(cols, rows), WithAltRef(window) and
WithFastEncode().  coding="raw" produces lossless CodingRaw key frames, the
encoder's default (encoder.go:134-143).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import native


@dataclass
class EncodedClip:
    container: bytes          # whole .ebiv file (header, records, index, footer)
    packets: list             # frame payloads in stream order (what Reader.NextPacket returns)
    recons: list              # encoder reconstruction per frame: (Y, Cb, Cr) packed, or []
    width: int
    height: int
    # token log (log_tokens=True): one Payload per coded payload in stream order
    payloads: list = field(default_factory=list)


@dataclass
class Payload:
    """A coded payload's token streams as the encoder produced them."""
    packet: int               # index of the packet the payload went into (a key frame's override shares it)
    inter: bool
    grid: tuple               # (tile_cols, tile_rows) after newTileGrid's clamp
    mbs: tuple                # (mbc, mbr)
    tiles: list               # per tile: uint16 array (tokens, 2) of (ctx, sym)
    freqs: np.ndarray         # (contexts, 16) frequency vectors, zero where none shipped


class Encoder:
    def __init__(self, width: int, height: int, *, qp: int = 20, gop: int = 1, tile_cols: int = 1,
                 tile_rows: int = 1, altref: int = 0, fast: bool = True, coding: str = "coded",
                 fps_num: int = 30, fps_den: int = 1, threads: int = 8, keep_recon: bool = False,
                 log_tokens: bool = False):
        self._L = native.encoder_lib()
        self.width, self.height = width, height
        self.cw, self.ch = (width + 1) // 2, (height + 1) // 2
        self._keep = keep_recon
        cod = {"raw": 0, "coded": 1}[coding]
        self._h = self._L.ebivenc_new(width, height, fps_num, fps_den, tile_cols, tile_rows, qp, gop, altref,
                                      cod, int(fast), threads, int(keep_recon))
        if not self._h:
            raise ValueError("invalid encoder configuration")
        if log_tokens:
            self._L.ebivenc_log_tokens(self._h, 1)

    def write(self, y: np.ndarray, cb: np.ndarray, cr: np.ndarray) -> None:
        y = np.ascontiguousarray(y, np.uint8)
        cb = np.ascontiguousarray(cb, np.uint8)
        cr = np.ascontiguousarray(cr, np.uint8)
        assert y.shape == (self.height, self.width) and cb.shape == (self.ch, self.cw) == cr.shape
        self._L.ebivenc_write(self._h, y.ctypes.data, cb.ctypes.data, cr.ctypes.data)

    def close(self) -> EncodedClip:
        L = self._L
        n = L.ebivenc_close(self._h)
        data = C.string_at(L.ebivenc_data(self._h), n)
        packets = []
        for i in range(L.ebivenc_packet_count(self._h)):
            ln = C.c_size_t()
            p = L.ebivenc_packet(self._h, i, C.byref(ln))
            packets.append(C.string_at(p, ln.value))
        recons = []
        ys, cs = self.width * self.height, self.cw * self.ch
        for i in range(L.ebivenc_recon_count(self._h)):
            raw = np.frombuffer(C.string_at(L.ebivenc_recon(self._h, i), ys + 2 * cs), np.uint8)
            recons.append((raw[:ys].reshape(self.height, self.width),
                           raw[ys:ys + cs].reshape(self.ch, self.cw),
                           raw[ys + cs:].reshape(self.ch, self.cw)))
        payloads = []
        info = (C.c_int32 * 6)()
        freqs = np.zeros((NUM_CONTEXTS, 16), np.uint32)
        for i in range(L.ebivenc_payload_count(self._h)):
            assert L.ebivenc_payload_info(self._h, i, info) == 0
            tiles = []
            for t in range(info[2] * info[3]):
                n = C.c_size_t()
                p = L.ebivenc_payload_tokens(self._h, i, t, C.byref(n))
                tiles.append(np.frombuffer(C.string_at(p, 4 * n.value), np.uint16).reshape(n.value, 2).copy()
                             if n.value else np.zeros((0, 2), np.uint16))
            assert L.ebivenc_payload_freqs(self._h, i, freqs.ctypes.data) == 0
            payloads.append(Payload(info[0], bool(info[1]), (info[2], info[3]), (info[4], info[5]), tiles,
                                    freqs.copy()))
        L.ebivenc_free(self._h)
        self._h = None
        return EncodedClip(data, packets, recons, self.width, self.height, payloads)


NUM_CONTEXTS = 72  # ctxTokenBase (12) + 45 luma + 15 chroma token contexts (model.go:124-149)


def encode_frames(frames, width, height, **kw) -> EncodedClip:
    """Encodes an iterable of (Y, Cb, Cr) packed frames."""
    enc = Encoder(width, height, **kw)
    for y, cb, cr in frames:
        enc.write(y, cb, cr)
    return enc.close()


def encode_blocks(levels: np.ndarray, plane: int, tx: int, sdh: bool = True):
    """Encoder's block grammar on hand-written levels (count, n, n).

    Returns (adjusted_levels, tables_blob, stream)."""
    L = native.encoder_lib()
    lv = np.ascontiguousarray(levels, np.int32).copy()
    count, n = lv.shape[0], lv.shape[1]
    tb = (C.c_uint8 * 8192)()
    st = (C.c_uint8 * (1 << 20))()
    tl, sl = C.c_size_t(8192), C.c_size_t(1 << 20)
    r = L.ebivenc_blocks(plane, tx, n, count, lv.ctypes.data, int(sdh), tb, C.byref(tl), st, C.byref(sl))
    assert r >= 0
    return lv, bytes(tb[:tl.value]), bytes(st[:sl.value])


def encode_tokens(ctxs, syms):
    L = native.encoder_lib()
    n = len(ctxs)
    tb = (C.c_uint8 * 8192)()
    st = (C.c_uint8 * (64 * n + 64))()
    tl, sl = C.c_size_t(8192), C.c_size_t(64 * n + 64)
    r = L.ebivenc_tokens((C.c_int * n)(*ctxs), (C.c_int * n)(*syms), n, tb, C.byref(tl), st, C.byref(sl))
    assert r == 0
    return bytes(tb[:tl.value]), bytes(st[:sl.value])


def normalize_freqs(counts):
    L = native.encoder_lib()
    n = len(counts)
    out = (C.c_uint32 * n)()
    ok = L.ebivenc_normalize((C.c_uint32 * n)(*counts), n, out)
    return list(out) if ok else None


def forward_dct(block: np.ndarray) -> np.ndarray:
    """forwardDCT (transform.go:44-79) of an n x n int32 block."""
    src = np.ascontiguousarray(block, np.int32)
    dst = np.zeros_like(src)
    assert native.encoder_lib().ebivenc_forward_dct(src.ctypes.data, dst.ctypes.data, src.shape[0]) == 0
    return dst


def quantize(coeffs: np.ndarray, qp: int) -> np.ndarray:
    """Dead-zone quantizer.quantize (quant.go:58-70)."""
    src = np.ascontiguousarray(coeffs, np.int32)
    dst = np.zeros_like(src)
    native.encoder_lib().ebivenc_quantize(qp, src.ctypes.data, dst.ctypes.data, src.shape[0])
    return dst
