"""The container reader in C++ (ebiv_reader_* in the C ABI; no GPU needed)
against the reference's container and format tests, restated
(ebiv/format_test.go:29-89, ebiv/container_test.go:49-270), and against the
Python restatement on every stream in the repository."""
import glob
import os
import struct

import pytest

from govid_amd import native
from govid_amd.container import (FILE_HEADER_SIZE, FOOTER_SIZE, FRAME_RECORD_SIZE, INDEX_ENTRY_SIZE,
                                  ErrBadMagic, ErrCorruptContainer, ErrUnsupportedContainer, ErrVersion,
                                  NativeReader, Reader)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STREAMS = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.ebiv")) +
                 glob.glob(os.path.join(ROOT, "bench_data", "*.ebiv")))
SECOND = 1_000_000_000


def _bytes(name):
    with open(os.path.join(ROOT, "tests", "golden", name), "rb") as f:
        return f.read()


@pytest.mark.parametrize("path", STREAMS, ids=os.path.basename)
def test_native_reader_equals_python_reader(path):
    data = open(path, "rb").read()
    a, b = Reader(data), NativeReader(data)
    assert a.header == b.header
    assert a.index == b.index and not b.index_scanned
    assert a.duration() == b.duration()
    assert a.packets() == b.packets()


def test_all_key_raw_stream_sequence_and_seek():
    """TestDemuxerPacketSequence / TestDemuxerSeek: timestamps, sticky EOF,
    every seek lands on its frame, past the end clamps, negative is 0."""
    data = _bytes("raw_48x32.ebiv")
    r = NativeReader(data)
    hdr = r.header
    n = r.frame_count()
    assert n >= 3 and all(e.keyframe for e in r.index)
    for i in range(n):
        p = r.next_packet()
        assert p is not None and p.keyframe and p.data
        assert p.timestamp == hdr.frame_time(i)
    assert r.next_packet() is None and r.next_packet() is None
    for target in (n - 1, 0, n // 2, 1):
        assert r.seek(hdr.frame_time(target)) == hdr.frame_time(target)
        assert r.next_packet().timestamp == hdr.frame_time(target)
    assert r.seek(3600 * SECOND) == hdr.frame_time(n - 1)
    assert r.seek(-SECOND) == 0
    r.close()
    assert r.next_packet() is None
    with pytest.raises(EOFError):
        r.seek(0)


def test_seek_goes_back_to_the_governing_key_frame():
    path = os.path.join(ROOT, "bench_data", "parallax_1080p.ebiv")
    py, r = Reader.open(path), NativeReader.open(path)
    keys = [i for i, e in enumerate(r.index) if e.keyframe]
    assert keys[0] == 0 and len(keys) >= 2
    for target in (0, 1, keys[1] - 1, keys[1], keys[1] + 5, r.frame_count() - 1):
        t = r.header.frame_time(target)
        want = max(k for k in keys if k <= target)
        assert r.seek(t) == r.header.frame_time(want) == py.seek(t)
        assert r.next_packet().data == py.next_packet().data


def test_frame_time_is_exact_for_ntsc_rates():
    """TestFrameTimeIsExactForNTSCRates (format_test.go:70-89)."""
    L = native.gpu_lib()
    ft = lambda i: L.ebiv_frame_time(30000, 1001, i)
    fi = lambda t: L.ebiv_frame_index_at(30000, 1001, t)
    assert ft(30000) == 1001 * SECOND
    for i in (0, 1, 29, 30, 1000, 30000):
        assert fi(ft(i)) == i
    assert fi(ft(10) - 1) == 9
    assert fi(-SECOND) == 0
    # the 30 fps example of the comment at format.go:257-262
    assert L.ebiv_frame_time(30, 1, 7) == 233333333 and L.ebiv_frame_index_at(30, 1, 233333333) == 7


def _header_variant(mutate):
    b = bytearray(_bytes("raw_48x32.ebiv"))
    mutate(b)
    return bytes(b)


@pytest.mark.parametrize("mutate,err", [
    (lambda b: b.__setitem__(0, ord("X")), ErrBadMagic),
    (lambda b: b.__setitem__(4, 5), ErrVersion),
    (lambda b: b.__setitem__(6, FILE_HEADER_SIZE + 1), ErrCorruptContainer),
    (lambda b: b.__setitem__(slice(8, 12), b"\0\0\0\0"), ErrCorruptContainer),
    (lambda b: b.__setitem__(slice(16, 20), b"\0\0\0\0"), ErrCorruptContainer),
    (lambda b: b.__setitem__(28, 1), ErrUnsupportedContainer),
    (lambda b: b.__setitem__(29, 1), ErrUnsupportedContainer),
], ids=["bad magic", "future version", "header size", "zero width", "zero frame rate", "chroma 4:2:2", "10-bit"])
def test_file_header_rejects(mutate, err):
    """TestFileHeaderRejects (format_test.go:29-68)."""
    data = _header_variant(mutate)
    with pytest.raises(err):
        NativeReader(data)
    with pytest.raises(err):
        Reader(data)


def test_rejects_corrupt_files():
    """TestDemuxerRejectsCorruptFiles (container_test.go:194-233)."""
    full = _bytes("cfg1_rects_320x180.ebiv")
    rec = bytearray(full[:FILE_HEADER_SIZE + FRAME_RECORD_SIZE])
    rec[FILE_HEADER_SIZE + 4:] = b"\xff" * 4
    nosync = bytearray(full[:-FOOTER_SIZE])
    nosync[FILE_HEADER_SIZE] ^= 0xFF
    cases = [(b"", ErrBadMagic), (b"X" + full[1:], ErrBadMagic), (full[:10], ErrBadMagic),
             (full[:FILE_HEADER_SIZE], ErrCorruptContainer), (bytes(rec), ErrCorruptContainer),
             (bytes(nosync), ErrCorruptContainer)]
    for data, err in cases:
        with pytest.raises(err):
            NativeReader(data)
        with pytest.raises(err):
            Reader(data)


def _payload_end(data):
    off = struct.unpack_from("<Q", data, len(data) - FOOTER_SIZE)[0]
    return off


@pytest.mark.parametrize("damage", ["no footer", "footer magic", "index checksum", "index bounds"])
def test_index_rebuilt_when_footer_missing_or_damaged(damage):
    """TestDemuxerRebuildsIndexWithoutFooter / TestDemuxerFallsBackWhenIndexIsCorrupt
    (container_test.go:163-192, 235-270): the record scan gives the same frames."""
    full = _bytes("cfg1_rects_320x180.ebiv")
    good = NativeReader(full)
    b = bytearray(full)
    if damage == "no footer":
        b = b[:_payload_end(full)]
    elif damage == "footer magic":
        b[-1] = ord("Z")
    elif damage == "index checksum":
        b[len(b) - FOOTER_SIZE - INDEX_ENTRY_SIZE] = (b[len(b) - FOOTER_SIZE - INDEX_ENTRY_SIZE] + 1) & 0xFF
    else:
        struct.pack_into("<Q", b, len(b) - FOOTER_SIZE, 2**64 - 8)  # an offset Go's Seek rejects
    r = NativeReader(bytes(b))
    assert r.index_scanned
    assert r.index == good.index
    assert r.packets() == good.packets() == Reader(bytes(b)).packets()


def test_trailing_partial_record_ends_the_scan():
    """scanFrames keeps everything before a cut record (reader.go:123-126)."""
    full = _bytes("cfg1_rects_320x180.ebiv")
    good = NativeReader(full)
    cut = full[:good.index[3].offset + FRAME_RECORD_SIZE + 5]
    r = NativeReader(cut)
    assert r.index_scanned and r.frame_count() == 3
    assert r.packets() == good.packets()[:3]
