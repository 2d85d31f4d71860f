"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE).

The oracle is a plain-C restatement of govid's Go EBIV decoder
(ebiv/codec.go et al.) and of the root package's YCbCr->RGBA conversion
(frame.go:8-67).  Only tests, the smoke check and the benchmark's CPU-baseline
leg may use it, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libebiv_oracle.so")
_lib = None

STATUS_NAMES = {0: "ok", 1: "corrupt", 2: "unsupported", 3: "nokeyframe", 4: "dimensions"}


class _Frame(C.Structure):
    _fields_ = [("y", C.c_void_p), ("cb", C.c_void_p), ("cr", C.c_void_p),
                ("y_stride", C.c_int), ("c_stride", C.c_int),
                ("width", C.c_int), ("height", C.c_int),
                ("cwidth", C.c_int), ("cheight", C.c_int)]


class _MBTrace(C.Structure):
    _fields_ = [("kind", C.c_int32), ("tx", C.c_int32), ("cmode", C.c_int32), ("refsel", C.c_int32),
                ("part", C.c_int32), ("cbp", C.c_int32), ("lmode", C.c_int32 * 16), ("mv", C.c_int16 * 8),
                ("ncoef", C.c_int32), ("coef", C.c_uint32 * 384)]


def build() -> str:
    """Compiles the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, u8p, i32p = C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_int32)
        L.orc_codec_new.restype = vp
        L.orc_codec_new.argtypes = [C.c_int]
        L.orc_codec_free.argtypes = [vp]
        L.orc_codec_decode.restype = C.c_int
        L.orc_codec_decode.argtypes = [vp, C.c_char_p, C.c_size_t, C.POINTER(_Frame)]
        L.orc_codec_flush.argtypes = [vp]
        L.orc_codec_last_tile_failure.restype = C.c_int
        L.orc_codec_last_tile_failure.argtypes = [vp]
        L.orc_codec_last_error.restype = C.c_char_p
        L.orc_codec_last_error.argtypes = [vp]
        L.orc_codec_set_threads.argtypes = [vp, C.c_int]
        L.orc_codec_symbols.restype = C.c_uint64
        L.orc_codec_symbols.argtypes = [vp]
        L.orc_codec_golden.restype = C.c_int
        L.orc_codec_golden.argtypes = [vp, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_codec_enable_trace.argtypes = [vp, C.c_int]
        L.orc_codec_trace.restype = C.POINTER(_MBTrace)
        L.orc_codec_trace.argtypes = [vp, C.POINTER(C.c_int)]
        L.orc_ycbcr_to_rgba.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, vp]
        L.orc_inverse_dct.restype = C.c_int
        L.orc_inverse_dct.argtypes = [i32p, i32p, C.c_int]
        L.orc_dct_coef.restype = C.c_int32
        L.orc_dct_coef.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_quant_step.restype = C.c_int32
        L.orc_quant_step.argtypes = [C.c_int]
        L.orc_zigzag.restype = C.c_int
        L.orc_zigzag.argtypes = [C.c_int, C.c_int]
        L.orc_predict_intra.restype = C.c_int
        L.orc_predict_intra.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int, C.c_int, i32p]
        L.orc_above_right_available.restype = C.c_int
        L.orc_above_right_available.argtypes = [C.c_int] * 7
        L.orc_mc_luma.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_int, i32p, C.c_int]
        L.orc_mc_chroma.argtypes = L.orc_mc_luma.argtypes
        L.orc_predict_mv.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(C.c_int16)]
        L.orc_rans_decode_symbols.restype = C.c_int
        L.orc_rans_decode_symbols.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                              C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
        L.orc_decode_blocks.restype = C.c_int
        L.orc_decode_blocks.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t, C.c_int,
                                        C.c_int, C.c_int, C.c_int, i32p]
        _lib = L
    return _lib


def _i32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class DecodeError(Exception):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class OracleCodec:
    """Codec (ebiv/codec.go) restated in C.  decode() returns packed planes."""

    def __init__(self, ring: int = 3, threads: int = 1):
        self._L = lib()
        self._h = self._L.orc_codec_new(ring)
        self._L.orc_codec_set_threads(self._h, threads)

    def close(self):
        if self._h:
            self._L.orc_codec_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_threads(self, n: int):
        self._L.orc_codec_set_threads(self._h, n)

    def decode_raw(self, data: bytes):
        """Decodes without copying the ring image out; returns (status, frame struct)."""
        fr = _Frame()
        st = self._L.orc_codec_decode(self._h, data, len(data), C.byref(fr))
        return st, fr

    def decode(self, data: bytes):
        """Returns (Y, Cb, Cr) packed uint8 arrays, or raises DecodeError."""
        st, fr = self.decode_raw(data)
        if st != 0:
            raise DecodeError(st, self._L.orc_codec_last_error(self._h).decode())
        return _planes(fr)

    def flush(self):
        self._L.orc_codec_flush(self._h)

    def last_tile_failure(self) -> int:
        """0: the last decode did not fail in a tile; 1: its own payload's
        tiles failed; 2: its golden override's tiles failed."""
        return int(self._L.orc_codec_last_tile_failure(self._h))

    def enable_trace(self, on: bool = True):
        self._L.orc_codec_enable_trace(self._h, int(on))

    def trace(self):
        """Per-macroblock trace of the last coded decode (list of dicts)."""
        n = C.c_int()
        p = self._L.orc_codec_trace(self._h, C.byref(n))
        out = []
        for i in range(n.value):
            t = p[i]
            out.append({"kind": t.kind, "tx": t.tx, "cmode": t.cmode, "refsel": t.refsel, "part": t.part,
                        "cbp": t.cbp, "lmode": list(t.lmode), "mv": list(t.mv),
                        "coef": sorted(t.coef[k] for k in range(t.ncoef))})
        return out

    def symbols(self) -> int:
        return int(self._L.orc_codec_symbols(self._h))

    def golden(self):
        y, cb, cr = C.c_void_p(), C.c_void_p(), C.c_void_p()
        w, h = C.c_int(), C.c_int()
        if self._L.orc_codec_golden(self._h, C.byref(y), C.byref(cb), C.byref(cr), C.byref(w), C.byref(h)):
            return None
        W, H = w.value, h.value
        gy = np.ctypeslib.as_array((C.c_uint8 * (W * H)).from_address(y.value)).reshape(H, W).copy()
        gcb = np.ctypeslib.as_array((C.c_uint8 * (W * H // 4)).from_address(cb.value)).reshape(H // 2, W // 2).copy()
        gcr = np.ctypeslib.as_array((C.c_uint8 * (W * H // 4)).from_address(cr.value)).reshape(H // 2, W // 2).copy()
        return gy, gcb, gcr


def _plane(ptr, stride, w, h):
    buf = (C.c_uint8 * (stride * h)).from_address(ptr)
    return np.ctypeslib.as_array(buf).reshape(h, stride)[:, :w].copy()


def _planes(fr: _Frame):
    return (_plane(fr.y, fr.y_stride, fr.width, fr.height),
            _plane(fr.cb, fr.c_stride, fr.cwidth, fr.cheight),
            _plane(fr.cr, fr.c_stride, fr.cwidth, fr.cheight))


def rgba(y: np.ndarray, cb: np.ndarray, cr: np.ndarray) -> np.ndarray:
    """convertYCbCr420ToRGBA (frame.go:8-67) on packed planes."""
    y = np.ascontiguousarray(y, np.uint8)
    cb = np.ascontiguousarray(cb, np.uint8)
    cr = np.ascontiguousarray(cr, np.uint8)
    h, w = y.shape
    out = np.empty((h, w, 4), np.uint8)
    lib().orc_ycbcr_to_rgba(y.ctypes.data, w, cb.ctypes.data, cr.ctypes.data, cb.shape[1], w, h, out.ctypes.data)
    return out


def inverse_dct(coeffs: np.ndarray) -> np.ndarray:
    n = coeffs.shape[0]
    src = np.ascontiguousarray(coeffs, np.int32)
    dst = np.zeros_like(src)
    assert lib().orc_inverse_dct(_i32(src), _i32(dst), n) == 0
    return dst


def dct_matrix(n: int) -> np.ndarray:
    L = lib()
    return np.array([[L.orc_dct_coef(n, k, i) for i in range(n)] for k in range(n)], np.int64)


def quant_step(qp: int) -> int:
    return int(lib().orc_quant_step(qp))


def zigzag(n: int) -> list:
    L = lib()
    return [L.orc_zigzag(n, i) for i in range(n * n)]


def predict_intra(plane: np.ndarray, x0, y0, n, mode, av_above, av_left, av_ar) -> np.ndarray:
    p = np.ascontiguousarray(plane, np.uint8)
    out = np.zeros(n * n, np.int32)
    h, w = p.shape
    rc = lib().orc_predict_intra(p.ctypes.data, w, w, h, x0, y0, n, mode, int(av_above), int(av_left),
                                 int(av_ar), _i32(out))
    assert rc == 0
    return out.reshape(n, n)


def above_right_available(x0, y0, n, tile_top, tile_right, mb_top, mb_right) -> bool:
    return bool(lib().orc_above_right_available(x0, y0, n, tile_top, tile_right, mb_top, mb_right))


def mc_luma(plane, sx, sy, fx, fy, pw, ph) -> np.ndarray:
    p = np.ascontiguousarray(plane, np.uint8)
    h, w = p.shape
    out = np.zeros(pw * ph, np.int32)
    lib().orc_mc_luma(p.ctypes.data, w, w, h, sx, sy, fx, fy, pw, ph, _i32(out), pw)
    return out.reshape(ph, pw)


def mc_chroma(plane, sx, sy, fx, fy, pw, ph) -> np.ndarray:
    p = np.ascontiguousarray(plane, np.uint8)
    h, w = p.shape
    out = np.zeros(pw * ph, np.int32)
    lib().orc_mc_chroma(p.ctypes.data, w, w, h, sx, sy, fx, fy, pw, ph, _i32(out), pw)
    return out.reshape(ph, pw)


def predict_mv(grid: np.ndarray, mbx, mby, x0, x1, y0, y1):
    g = np.ascontiguousarray(grid, np.int16)  # (rows, cols, 2)
    out = (C.c_int16 * 2)()
    lib().orc_predict_mv(g.ctypes.data, g.shape[1], mbx, mby, x0, x1, y0, y1, out)
    return int(out[0]), int(out[1])


def rans_decode_symbols(tables: bytes, stream: bytes, ctxs):
    n = len(ctxs)
    cs = (C.c_int * n)(*ctxs)
    out = (C.c_int * n)()
    st = lib().orc_rans_decode_symbols(tables, len(tables), stream, len(stream), cs, n, out)
    return st, list(out)


def decode_blocks(tables: bytes, stream: bytes, plane: int, tx: int, n: int, count: int):
    out = np.zeros(count * n * n, np.int32)
    st = lib().orc_decode_blocks(tables, len(tables), stream, len(stream), plane, tx, n, count, _i32(out))
    return st, out.reshape(count, n, n)
