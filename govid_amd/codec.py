"""Host mirror of govid's Codec interface over the MI355X decoder.

govid.Codec (govid.go:41-44) is ``Decode(Packet) (*Frame, error)`` plus
``Flush()``; ebiv.NewCodec(WithFrameRing(n)) builds one (ebiv/codec.go:45-64).
This module keeps those names and meanings: ``Codec.decode(Packet) -> Frame``
raises the Go sentinel errors as exceptions, ``flush()`` resets only the ring
position, and a returned Frame aliases a ring buffer that stays valid for
``ring`` further decodes (ebiv/codec.go:15-22).

Decoding always runs on the GPU through libebivgpu.so; there is no other path.
"""
from __future__ import annotations

import ctypes as C
import struct
from dataclasses import dataclass

import numpy as np

from . import capi, native


class EbivError(Exception):
    """Base of the ebiv sentinel errors (ebiv/format.go:110-119)."""


class ErrCorrupt(EbivError):
    pass


class ErrUnsupported(EbivError):
    pass


class ErrNoKeyframe(EbivError):
    pass


class ErrDimensions(EbivError):
    pass


class DeviceError(RuntimeError):
    pass


_ERRORS = {capi.ERR_CORRUPT: ErrCorrupt, capi.ERR_UNSUPPORTED: ErrUnsupported,
           capi.ERR_NOKEYFRAME: ErrNoKeyframe, capi.ERR_DIMENSIONS: ErrDimensions}


@dataclass
class Packet:
    """govid.Packet (govid.go:9-13); timestamp in nanoseconds."""
    data: bytes
    timestamp: int = 0
    keyframe: bool = False


class Frame:
    """govid.Frame (govid.go:16-22) with YCbCr planes as numpy views.

    Host frames view pinned ring memory; device frames carry HBM pointers and
    are read back with ``planes()``.
    """

    def __init__(self, codec, raw: capi.Frame):
        self._codec = codec
        self._raw = raw
        self.width = raw.width
        self.height = raw.height
        self.timestamp = raw.timestamp_ns
        self.on_device = bool(raw.on_device)

    def _host_plane(self, ptr, stride, w, h):
        buf = (C.c_uint8 * (stride * h)).from_address(ptr)
        return np.ctypeslib.as_array(buf).reshape(h, stride)[:, :w]

    def _device_plane(self, ptr, stride, w, h):
        out = np.empty((h, stride), np.uint8)
        rc = self._codec._L.ebiv_memcpy_d2h(out.ctypes.data, ptr, stride * h)
        if rc:
            raise DeviceError("plane download failed")
        return out[:, :w]

    @property
    def Y(self):
        r = self._raw
        f = self._device_plane if self.on_device else self._host_plane
        return f(r.y, r.y_stride, r.width, r.height)

    @property
    def Cb(self):
        r = self._raw
        f = self._device_plane if self.on_device else self._host_plane
        return f(r.cb, r.c_stride, r.cwidth, r.cheight)

    @property
    def Cr(self):
        r = self._raw
        f = self._device_plane if self.on_device else self._host_plane
        return f(r.cr, r.c_stride, r.cwidth, r.cheight)

    def planes(self):
        """Packed copies of (Y, Cb, Cr)."""
        return self.Y.copy(), self.Cb.copy(), self.Cr.copy()

    def rgba(self) -> np.ndarray:
        """Frame.RGBA (frame.go:85-91), converted on the GPU, as an (h, w, 4) array."""
        out = np.empty((self.height, self.width, 4), np.uint8)
        rc = self._codec._rgba(self._raw, out.ctypes.data, 0)
        if rc:
            raise DeviceError(f"rgba conversion failed ({rc})")
        return out

    def rgba_device(self, dst_ptr: int) -> None:
        """Writes RGBA into device memory at dst_ptr (width*height*4 bytes)."""
        rc = self._codec._rgba(self._raw, dst_ptr, 1)
        if rc:
            raise DeviceError(f"rgba conversion failed ({rc})")


class Codec:
    """ebiv.Codec on the GPU: ``Codec(ring=3).decode(Packet) -> Frame``."""

    def __init__(self, ring: int = 3, device: int = 0):
        self._L = native.gpu_lib()
        h = C.c_void_p()
        rc = self._L.ebiv_codec_create(ring, device, C.byref(h))
        if rc:
            raise DeviceError(f"ebiv_codec_create failed ({rc})")
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            self._L.ebiv_codec_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc == capi.OK:
            return
        msg = self._L.ebiv_codec_last_error(self._h).decode()
        if rc in _ERRORS:
            raise _ERRORS[rc](msg)
        raise DeviceError(f"{msg} ({rc})")

    def decode(self, pkt) -> Frame:
        """Codec.Decode (ebiv/codec.go:70-111); pkt is a Packet or bytes."""
        if isinstance(pkt, (bytes, bytearray)):
            pkt = Packet(bytes(pkt))
        fr = capi.Frame()
        rc = self._L.ebiv_codec_decode(self._h, bytes(pkt.data), len(pkt.data), int(pkt.timestamp), C.byref(fr))
        self._check(rc)
        return Frame(self, fr)

    def decode_device(self, host_data: bytes, dev_ptr: int, timestamp: int = 0) -> Frame:
        """Decode of a packet already in HBM; the frame stays in HBM."""
        fr = capi.Frame()
        rc = self._L.ebiv_codec_decode_device(self._h, host_data, dev_ptr, len(host_data), int(timestamp), C.byref(fr))
        self._check(rc)
        return Frame(self, fr)

    def flush(self):
        """Codec.Flush (ebiv/codec.go:115-117)."""
        self._L.ebiv_codec_flush(self._h)

    def set_timing(self, enabled: bool) -> None:
        """Brackets every kernel launch with HIP events (measurement only)."""
        self._L.ebiv_codec_set_timing(self._h, int(enabled))

    def kernel_times(self) -> dict:
        t = capi.KernelTimes()
        self._L.ebiv_codec_kernel_times(self._h, C.byref(t))
        return {f: getattr(t, f) for f, _ in capi.KernelTimes._fields_}

    def debug_dump(self, slot: int = 0):
        """Entropy-kernel output of the last coded decode: (records, coefs, status)
        as numpy structured arrays (diagnostics)."""
        rec_dt = np.dtype([("kind", "u1"), ("part", "u1"), ("cbp", "u1"), ("tx", "u1"), ("refsel", "u1"),
                           ("cmode", "u1"), ("pad0", "u1"), ("pad1", "u1"), ("wmask", "<u4"),
                           ("coef_off", "<u4"), ("coef_count", "<u4"), ("mv", "<i2", (4, 2)),
                           ("lmode", "<u4", (2,)), ("pad2", "<u4")])
        st_dt = np.dtype([("err", "<i4"), ("mbs_done", "<u4"), ("nsym", "<u4"), ("n_intra", "<u4")])
        n = self._L.ebiv_codec_debug_dump(self._h, slot, None, 0, None, 0, None, 0)
        if n < 0:
            raise DeviceError("debug dump failed")
        recs = np.zeros(n, rec_dt)
        coefs = np.zeros(n * 384, np.uint32)
        status = np.zeros(4096, st_dt)
        self._L.ebiv_codec_debug_dump(self._h, slot, recs.ctypes.data, recs.nbytes, coefs.ctypes.data, coefs.size,
                                      status.ctypes.data, status.nbytes)
        return recs, coefs, status

    def stats(self) -> dict:
        s = capi.Stats()
        self._L.ebiv_codec_stats(self._h, C.byref(s))
        return {f: getattr(s, f) for f, _ in capi.Stats._fields_}

    def _rgba(self, raw, dst, on_device):
        return self._L.ebiv_frame_rgba(self._h, C.byref(raw), dst, on_device)


class DeviceBuffer:
    """A block of HBM holding bytes (packets kept resident for decode_device)."""

    def __init__(self, data: bytes, device: int = 0):
        self.ptr = 0
        self._L = native.gpu_lib()
        p = C.c_void_p()
        rc = self._L.ebiv_device_alloc(device, max(len(data), 1), C.byref(p))
        if rc:
            raise DeviceError("device allocation failed")
        self.ptr = p.value
        self.size = len(data)
        if data:
            src = C.create_string_buffer(bytes(data), len(data))
            if self._L.ebiv_memcpy_h2d(self.ptr, src, len(data)):
                raise DeviceError("upload failed")

    def close(self):
        if self.ptr:
            self._L.ebiv_device_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        self.close()


def device_count() -> int:
    return int(native.gpu_lib().ebiv_device_count())


class Completion:
    """One decoded (or failed) packet of a batch step.  ``frame`` (None on
    failure) and ``error`` are built on first access."""

    __slots__ = ("session", "tag", "status", "surface", "_batch", "_raw", "_msg", "_frame")

    def __init__(self, session: int, tag: int, status: int, batch, raw: bytes | None, msg: bytes = b"",
                 surface: int = -1):
        self.session, self.tag, self.status, self.surface = session, tag, status, surface
        self._batch, self._raw, self._msg, self._frame = batch, raw, msg, None

    @property
    def frame(self) -> Frame | None:
        if self.status or self._raw is None:
            return None
        if self._frame is None:
            self._frame = Frame(self._batch, capi.Frame.from_buffer_copy(self._raw))
        return self._frame

    @property
    def error(self) -> EbivError | None:
        if not self.status:
            return None
        return _ERRORS.get(self.status, DeviceError)(self._msg.split(b"\0", 1)[0].decode(errors="replace"))


class Batch:
    """Many ebiv.Codec sessions decoded together on one GPU.

    ``open()`` makes a session (one stream's codec state), ``submit()`` queues
    its packets, ``step()`` decodes up to ``window`` queued packets of every
    session in one entropy launch plus one reconstruction launch per frame
    position, and returns a Completion per packet in submission order per
    session.  Frames are HBM planes valid until ``ring`` further frames of
    their session complete.
    """

    def __init__(self, device: int = 0):
        self._L = native.gpu_lib()
        h = C.c_void_p()
        rc = self._L.ebiv_batch_create(device, C.byref(h))
        if rc:
            raise DeviceError(f"ebiv_batch_create failed ({rc})")
        self._h = h
        self.device = device
        self._out = (capi.Completion * 0)()

    def close(self):
        if getattr(self, "_h", None):
            self._L.ebiv_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def open(self, ring: int = 3, window: int = 1) -> int:
        sid = C.c_int()
        rc = self._L.ebiv_batch_open(self._h, ring, window, C.byref(sid))
        if rc:
            raise DeviceError(f"ebiv_batch_open failed ({rc})")
        return sid.value

    def close_session(self, session: int) -> None:
        if self._L.ebiv_batch_close(self._h, session):
            raise DeviceError("unknown session")

    def submit(self, session: int, pkt, dev_ptr: int | None = None, tag: int = 0) -> None:
        if isinstance(pkt, (bytes, bytearray)):
            pkt = Packet(bytes(pkt))
        data = bytes(pkt.data)
        rc = self._L.ebiv_batch_submit(self._h, session, data, dev_ptr, len(data), int(pkt.timestamp), tag)
        if rc:
            raise DeviceError(f"ebiv_batch_submit failed ({rc})")

    def submit_many(self, session: int, packets, dev_ptrs=None, tags=None, timestamps=None) -> None:
        """submit() of a list of packets (bytes) in one C call."""
        n = len(packets)
        data = [bytes(p) for p in packets]
        hp = (C.c_void_p * n)(*[C.cast(C.c_char_p(d), C.c_void_p).value for d in data])
        dp = (C.c_void_p * n)(*dev_ptrs) if dev_ptrs is not None else None
        ln = (C.c_size_t * n)(*[len(d) for d in data])
        ts = (C.c_int64 * n)(*timestamps) if timestamps is not None else None
        tg = (C.c_uint64 * n)(*tags) if tags is not None else None
        rc = self._L.ebiv_batch_submit_many(self._h, session, n, hp, dp, ln, ts, tg)
        if rc:
            raise DeviceError(f"ebiv_batch_submit_many failed ({rc})")

    def flush(self, session: int) -> None:
        """Codec.Flush for one session; its queued packets are dropped."""
        if self._L.ebiv_batch_flush(self._h, session):
            raise DeviceError("unknown session")

    def pending(self, session: int = 0) -> int:
        return int(self._L.ebiv_batch_pending(self._h, session))

    def step(self, max_out: int = 4096) -> list:
        if len(self._out) < max_out:
            self._out = (capi.Completion * max_out)()
        n = C.c_int()
        rc = self._L.ebiv_batch_step(self._h, self._out, max_out, C.byref(n))
        if rc:
            raise DeviceError(f"{self._L.ebiv_batch_last_error(self._h).decode()} ({rc})")
        # one copy of the completions, then plain slices (the output array is
        # reused by the next step)
        size, fo, fs = C.sizeof(capi.Completion), capi.Completion.frame.offset, C.sizeof(capi.Frame)
        mo, ms = capi.Completion.message.offset, capi.Completion.message.size
        so = capi.Completion.surface.offset
        raw = C.string_at(C.addressof(self._out), size * n.value)
        head = struct.Struct("<iiQ")
        surf = struct.Struct("<i")
        res = []
        for i in range(n.value):
            o = i * size
            session, status, tag = head.unpack_from(raw, o)
            if status == capi.OK:
                res.append(Completion(session, tag, 0, self, raw[o + fo:o + fo + fs], b"",
                                      surf.unpack_from(raw, o + so)[0]))
            else:
                res.append(Completion(session, tag, status, self, None, raw[o + mo:o + mo + ms]))
        return res

    def set_surfaces(self, session: int, ptrs, pitch: int) -> None:
        """Presentation: the session's frames are also written as RGBA into
        these device surfaces (rows of `pitch` bytes), in turn; a
        completion's ``surface`` names the one its frame went to."""
        arr = (C.c_void_p * max(len(ptrs), 1))(*ptrs)
        if self._L.ebiv_batch_set_surfaces(self._h, session, arr, len(ptrs), pitch):
            raise DeviceError("ebiv_batch_set_surfaces failed")

    def stats(self, session: int = 0) -> dict:
        s = capi.Stats()
        self._L.ebiv_batch_stats(self._h, session, C.byref(s))
        return {f: getattr(s, f) for f, _ in capi.Stats._fields_}

    def set_timing(self, enabled: bool) -> None:
        self._L.ebiv_batch_set_timing(self._h, int(enabled))

    def kernel_times(self) -> dict:
        t = capi.KernelTimes()
        self._L.ebiv_batch_kernel_times(self._h, C.byref(t))
        return {f: getattr(t, f) for f, _ in capi.KernelTimes._fields_}

    def set_schedule(self, **kw) -> None:
        """Launch shapes of the reconstruction (fields of capi.Schedule:
        intra wavefronts per tile, recon_lanes streams; omitted ones keep
        their value).  Results do not depend on them."""
        s = capi.Schedule(**{k: int(v) for k, v in kw.items()})
        if self._L.ebiv_batch_set_schedule(self._h, C.byref(s)):
            raise ValueError("bad schedule")

    def schedule(self) -> dict:
        s = capi.Schedule()
        self._L.ebiv_batch_get_schedule(self._h, C.byref(s))
        return {f: getattr(s, f) for f, _ in capi.Schedule._fields_}

    def _rgba(self, raw, dst, on_device):
        return self._L.ebiv_batch_frame_rgba(self._h, C.byref(raw), dst, on_device)


def session_bytes(width: int, height: int, ring: int = 3, window: int = 1) -> int:
    """HBM one batch session holds at this geometry."""
    return int(native.gpu_lib().ebiv_session_bytes(width, height, ring, window))
