"""EBIV container reader — the host side of the decode path.

Restates ebiv.NewDemuxer / Reader (ebiv/reader.go:32-258) and the file header
(ebiv/format.go:151-247): fixed 48-byte header, self-delimiting frame records
(8-byte prefix with the 0xEB17 sync word), an index of 16-byte entries and a
20-byte footer with a CRC-32 of the index.  A missing or corrupt footer falls
back to scanning the records.  Packets stay on the host; the decoder takes
them through the C ABI.
"""
from __future__ import annotations

import struct
import zlib
from dataclasses import dataclass

from .codec import Packet

FILE_HEADER_SIZE = 48
FRAME_RECORD_SIZE = 8
INDEX_ENTRY_SIZE = 16
FOOTER_SIZE = 20
VERSION = 4
MAX_FRAME_SIZE = 256 << 20
FRAME_SYNC = 0xEB17


class ContainerError(Exception):
    pass


class ErrBadMagic(ContainerError):
    pass


class ErrVersion(ContainerError):
    pass


class ErrCorruptContainer(ContainerError):
    pass


class ErrUnsupportedContainer(ContainerError):
    pass


@dataclass
class FileHeader:
    version: int
    width: int
    height: int
    fps_num: int
    fps_den: int
    chroma_format: int
    bit_depth: int
    color_range: int
    color_primaries: int
    tile_cols: int
    tile_rows: int
    flags: int

    def frame_time(self, i: int) -> int:
        """frameTime (format.go:250-253), nanoseconds."""
        return (i * self.fps_den * 1_000_000_000) // self.fps_num

    def frame_index_at(self, t: int) -> int:
        """frameIndexAt (format.go:263-275)."""
        if t <= 0:
            return 0
        i = (t * self.fps_num) // (self.fps_den * 1_000_000_000)
        while self.frame_time(i + 1) <= t:
            i += 1
        while i > 0 and self.frame_time(i) > t:
            i -= 1
        return i


@dataclass
class IndexEntry:
    offset: int
    size: int
    keyframe: bool


def parse_file_header(b: bytes) -> FileHeader:
    """FileHeader.unmarshal (format.go:208-247)."""
    if len(b) < FILE_HEADER_SIZE:
        raise ErrCorruptContainer("file header truncated")
    if b[0:4] != b"EBIV":
        raise ErrBadMagic("not an EBIV file")
    version, size = struct.unpack_from("<HH", b, 4)
    if version != VERSION:
        raise ErrVersion(f"version {version}, want {VERSION}")
    if size != FILE_HEADER_SIZE:
        raise ErrCorruptContainer(f"header size {size}")
    w, h, fn, fd = struct.unpack_from("<IIII", b, 8)
    if w == 0 or w > 65535 or h == 0 or h > 65535:
        raise ErrCorruptContainer(f"dimensions {w}x{h} out of range")
    if fn == 0 or fd == 0:
        raise ErrCorruptContainer(f"frame rate {fn}/{fd}")
    chroma, depth, crange, cprim = b[28], b[29], b[30], b[31]
    if chroma != 0:
        raise ErrUnsupportedContainer(f"chroma format {chroma}")
    if depth != 0:
        raise ErrUnsupportedContainer(f"bit depth code {depth}")
    tc, tr, flags = struct.unpack_from("<HHI", b, 32)
    return FileHeader(version, w, h, fn, fd, chroma, depth, crange, cprim, tc, tr, flags)


def _validate_entry(e: IndexEntry, payload_end: int):
    if e.offset < FILE_HEADER_SIZE or e.size == 0 or e.size > MAX_FRAME_SIZE:
        raise ErrCorruptContainer(f"index entry offset {e.offset} size {e.size}")
    if e.offset + FRAME_RECORD_SIZE + e.size > payload_end:
        raise ErrCorruptContainer(f"index entry at {e.offset} runs past the payload")


def read_index(data: bytes):
    """readIndex (reader.go:65-111)."""
    size = len(data)
    if size < FILE_HEADER_SIZE + FOOTER_SIZE:
        raise ErrCorruptContainer("file too short to hold a footer")
    footer = data[size - FOOTER_SIZE:]
    if footer[16:20] != b"EBIX":
        raise ErrCorruptContainer("missing footer magic")
    offset, count, want = struct.unpack_from("<QII", footer, 0)
    end = offset + count * INDEX_ENTRY_SIZE
    if offset < FILE_HEADER_SIZE or end != size - FOOTER_SIZE:
        raise ErrCorruptContainer("index bounds do not fit the file")
    buf = data[offset:end]
    if zlib.crc32(buf) & 0xFFFFFFFF != want:
        raise ErrCorruptContainer("index checksum mismatch")
    index = []
    for i in range(count):
        off, sz, flags = struct.unpack_from("<QII", buf, i * INDEX_ENTRY_SIZE)
        e = IndexEntry(off, sz, bool(flags & 1))
        _validate_entry(e, offset)
        index.append(e)
    if index and not index[0].keyframe:
        raise ErrCorruptContainer("first frame is not a key frame")
    return index


def scan_frames(data: bytes):
    """scanFrames (reader.go:126-160): rebuild the index from the records."""
    size = len(data)
    index = []
    off = FILE_HEADER_SIZE
    while off + FRAME_RECORD_SIZE <= size:
        sync, flags, payload = struct.unpack_from("<HHI", data, off)
        if sync != FRAME_SYNC or payload == 0 or payload > MAX_FRAME_SIZE:
            break
        end = off + FRAME_RECORD_SIZE + payload
        if end > size:
            break
        index.append(IndexEntry(off, payload, bool(flags & 1)))
        off = end
    if not index:
        raise ErrCorruptContainer("no frames found")
    if not index[0].keyframe:
        raise ErrCorruptContainer("first frame is not a key frame")
    return index


class Reader:
    """ebiv.Reader over an in-memory container (reader.go:20-258)."""

    def __init__(self, data: bytes):
        if len(data) < FILE_HEADER_SIZE:
            raise ErrBadMagic("not an EBIV file")
        self.data = bytes(data)
        self.header = parse_file_header(self.data)
        try:
            self.index = read_index(self.data)
        except ContainerError:
            self.index = scan_frames(self.data)
        self.next = 0
        self.closed = False

    @classmethod
    def open(cls, path: str) -> "Reader":
        with open(path, "rb") as f:
            return cls(f.read())

    def frame_count(self) -> int:
        return len(self.index)

    def duration(self) -> int:
        return self.header.frame_time(len(self.index))

    def payload(self, i: int) -> bytes:
        e = self.index[i]
        return self.data[e.offset + FRAME_RECORD_SIZE:e.offset + FRAME_RECORD_SIZE + e.size]

    def next_packet(self):
        """NextPacket (reader.go:189-217); returns None at EOF."""
        if self.closed or self.next >= len(self.index):
            return None
        i = self.next
        self.next += 1
        return Packet(self.payload(i), self.header.frame_time(i), self.index[i].keyframe)

    def keyframe_before(self, i: int) -> int:
        if i >= len(self.index):
            i = len(self.index) - 1
        while i > 0 and not self.index[i].keyframe:
            i -= 1
        return i

    def seek(self, t: int) -> int:
        """Seek (reader.go:222-238): to the last key frame at or before t."""
        target = min(self.header.frame_index_at(t), len(self.index) - 1)
        target = self.keyframe_before(target)
        self.next = target
        return self.header.frame_time(target)

    def packets(self):
        return [self.payload(i) for i in range(len(self.index))]

    def close(self):
        self.closed = True


class NativeReader:
    """The same reader in C++ behind the C ABI (ebiv_reader_*,
    govid_amd/csrc/ebiv_container.cpp), with Reader's interface."""

    _ERRORS = {1: ErrCorruptContainer, 2: ErrUnsupportedContainer, 5: ErrBadMagic, 6: ErrVersion}

    def __init__(self, data: bytes):
        import ctypes as C

        from . import capi, native

        self._C, self._capi = C, capi
        self._L = native.gpu_lib()
        h = C.c_void_p()
        msg = C.create_string_buffer(256)
        rc = self._L.ebiv_reader_open(bytes(data), len(data), C.byref(h), msg, 256)
        if rc:
            raise self._ERRORS.get(rc, ContainerError)(msg.value.decode(errors="replace"))
        self._h = h
        fh = capi.FileHeader()
        self._L.ebiv_reader_header(h, C.byref(fh))
        self.header = FileHeader(fh.version, fh.width, fh.height, fh.fps_num, fh.fps_den, fh.chroma_format,
                                 fh.bit_depth, fh.color_range, fh.color_primaries, fh.tile_cols, fh.tile_rows,
                                 fh.flags)
        n = C.c_size_t()
        self._L.ebiv_reader_index(h, None, 0, C.byref(n))
        arr = (capi.IndexEntry * max(n.value, 1))()
        self._L.ebiv_reader_index(h, arr, n.value, C.byref(n))
        self.index = [IndexEntry(e.offset, e.size, bool(e.keyframe)) for e in arr[:n.value]]
        self.index_scanned = bool(self._L.ebiv_reader_index_scanned(h))

    @classmethod
    def open(cls, path: str) -> "NativeReader":
        with open(path, "rb") as f:
            return cls(f.read())

    def frame_count(self) -> int:
        return int(self._L.ebiv_reader_frame_count(self._h))

    def duration(self) -> int:
        return int(self._L.ebiv_reader_duration(self._h))

    def next_packet(self):
        """NextPacket; None at EOF."""
        C = self._C
        p, n, ts, key = C.c_void_p(), C.c_size_t(), C.c_int64(), C.c_int32()
        rc = self._L.ebiv_reader_next_packet(self._h, C.byref(p), C.byref(n), C.byref(ts), C.byref(key))
        if rc == self._capi.EOF:
            return None
        if rc:
            raise ErrCorruptContainer("packet outside the file")
        return Packet(C.string_at(p.value, n.value), ts.value, bool(key.value))

    def seek(self, t: int) -> int:
        C = self._C
        out = C.c_int64()
        rc = self._L.ebiv_reader_seek(self._h, int(t), C.byref(out))
        if rc == self._capi.EOF:
            raise EOFError("reader closed or empty")
        return out.value

    def packets(self):
        out = []
        self.seek(0)
        while True:
            p = self.next_packet()
            if p is None:
                break
            out.append(p.data)
        return out

    def close(self):
        self._L.ebiv_reader_close(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.ebiv_reader_destroy(h)
            self._h = None
