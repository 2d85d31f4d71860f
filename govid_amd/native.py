"""Loader for the package's native libraries (built in-tree under govid_amd/lib).

libebivgpu.so  — the HIP decoder behind the C ABI in include/ebiv_gpu.h.
libebivenc.so  — the host C++ stream encoder.

The GPU library is required: if it is missing or cannot be loaded the product
path raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
GPU_LIB = os.path.join(LIB_DIR, "libebivgpu.so")
ENC_LIB = os.path.join(LIB_DIR, "libebivenc.so")

_enc = None
_gpu = None


class NativeLibraryMissing(RuntimeError):
    pass


def _load(path: str):
    if not os.path.exists(path):
        raise NativeLibraryMissing(f"{path} is not built; run __graft_entry__.build()")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


def encoder_lib():
    global _enc
    if _enc is None:
        L = _load(ENC_LIB)
        vp, sz = C.c_void_p, C.c_size_t
        L.ebivenc_new.restype = vp
        L.ebivenc_new.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ebivenc_write.restype = C.c_int
        L.ebivenc_write.argtypes = [vp, vp, vp, vp]
        L.ebivenc_close.restype = sz
        L.ebivenc_close.argtypes = [vp]
        L.ebivenc_data.restype = vp
        L.ebivenc_data.argtypes = [vp]
        L.ebivenc_packet_count.restype = sz
        L.ebivenc_packet_count.argtypes = [vp]
        L.ebivenc_packet.restype = vp
        L.ebivenc_packet.argtypes = [vp, sz, C.POINTER(sz)]
        L.ebivenc_recon_count.restype = sz
        L.ebivenc_recon_count.argtypes = [vp]
        L.ebivenc_recon.restype = vp
        L.ebivenc_recon.argtypes = [vp, sz]
        L.ebivenc_free.argtypes = [vp]
        L.ebivenc_log_tokens.argtypes = [vp, C.c_int]
        L.ebivenc_payload_count.restype = sz
        L.ebivenc_payload_count.argtypes = [vp]
        L.ebivenc_payload_info.restype = C.c_int
        L.ebivenc_payload_info.argtypes = [vp, sz, vp]
        L.ebivenc_payload_tokens.restype = vp
        L.ebivenc_payload_tokens.argtypes = [vp, sz, C.c_int, C.POINTER(sz)]
        L.ebivenc_payload_freqs.restype = C.c_int
        L.ebivenc_payload_freqs.argtypes = [vp, sz, vp]
        L.ebivenc_blocks.restype = C.c_int
        L.ebivenc_blocks.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.POINTER(sz), vp,
                                     C.POINTER(sz)]
        L.ebivenc_tokens.restype = C.c_int
        L.ebivenc_tokens.argtypes = [vp, vp, C.c_int, vp, C.POINTER(sz), vp, C.POINTER(sz)]
        L.ebivenc_forward_dct.restype = C.c_int
        L.ebivenc_forward_dct.argtypes = [vp, vp, C.c_int]
        L.ebivenc_quantize.restype = C.c_int
        L.ebivenc_quantize.argtypes = [C.c_int, vp, vp, C.c_int]
        L.ebivenc_normalize.restype = C.c_int
        L.ebivenc_normalize.argtypes = [vp, C.c_int, vp]
        _enc = L
    return _enc


def gpu_lib():
    """The HIP decoder library; raises NativeLibraryMissing when absent."""
    global _gpu
    if _gpu is None:
        from . import capi  # noqa: F401  (declares prototypes)
        _gpu = capi.bind(_load(GPU_LIB))
    return _gpu
