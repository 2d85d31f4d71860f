"""The C ABI without a GPU: the library loads, exports every entry point the
header declares, and its host-only entry points behave."""
import ctypes as C
import os
import re

from govid_amd import capi, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ebiv_gpu.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(ebiv_[a-z0-9_]+)\s*\(", text))


def test_header_matches_binding_table():
    assert _declared() == set(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(native.GPU_LIB)
    missing = [n for n in sorted(_declared()) if not hasattr(L, n)]
    assert not missing, f"libebivgpu.so lacks {missing}"


def test_no_oracle_in_product_library():
    """The product library links nothing from the oracle."""
    data = open(native.GPU_LIB, "rb").read()
    assert b"orc_codec" not in data and b"ebiv_oracle" not in data


def test_session_bytes_host_only():
    L = native.gpu_lib()
    one = L.ebiv_session_bytes(1920, 1080, 3, 1)
    eight = L.ebiv_session_bytes(1920, 1080, 3, 8)
    # 1920x1088 padded luma + two 960x544 chroma planes per frame buffer
    frame = 1920 * 1088 + 2 * 960 * 544
    assert one >= 6 * frame and eight > one
    assert L.ebiv_session_bytes(0, 10, 3, 1) == 0


def test_null_handles_are_argument_errors():
    L = native.gpu_lib()
    assert L.ebiv_codec_create(3, 0, None) == capi.ERR_ARGUMENT
    assert L.ebiv_batch_create(0, None) == capi.ERR_ARGUMENT
    n = C.c_int()
    assert L.ebiv_batch_step(None, None, 0, C.byref(n)) == capi.ERR_ARGUMENT
    assert L.ebiv_codec_last_error(None) == b"null codec"
