"""ctypes prototypes of the C ABI in include/ebiv_gpu.h (libebivgpu.so)."""
from __future__ import annotations

import ctypes as C

OK = 0
ERR_CORRUPT = 1
ERR_UNSUPPORTED = 2
ERR_NOKEYFRAME = 3
ERR_DIMENSIONS = 4
ERR_BAD_MAGIC = 5
ERR_VERSION = 6
EOF = 7
ERR_ARGUMENT = 16
ERR_DEVICE = 17

# Every symbol include/ebiv_gpu.h declares; tests check the library exports them.
EXPORTS = (
    "ebiv_codec_create", "ebiv_codec_decode", "ebiv_codec_decode_device", "ebiv_codec_flush",
    "ebiv_codec_destroy", "ebiv_codec_last_error", "ebiv_codec_stats", "ebiv_frame_rgba", "ebiv_frame_rgba_noalloc",
    "ebiv_device_alloc", "ebiv_device_free", "ebiv_memcpy_h2d", "ebiv_memcpy_d2h", "ebiv_device_count",
    "ebiv_codec_set_timing", "ebiv_codec_kernel_times", "ebiv_codec_debug_dump",
    "ebiv_batch_create", "ebiv_batch_destroy", "ebiv_batch_last_error", "ebiv_batch_open", "ebiv_batch_close",
    "ebiv_batch_submit", "ebiv_batch_flush", "ebiv_batch_pending", "ebiv_batch_step", "ebiv_batch_stats",
    "ebiv_batch_set_timing", "ebiv_batch_kernel_times", "ebiv_batch_frame_rgba", "ebiv_session_bytes",
    "ebiv_batch_set_schedule", "ebiv_batch_get_schedule", "ebiv_batch_submit_many", "ebiv_batch_set_surfaces",
    "ebiv_reader_open", "ebiv_reader_destroy", "ebiv_reader_header", "ebiv_reader_frame_count", "ebiv_reader_index",
    "ebiv_reader_index_scanned", "ebiv_reader_duration", "ebiv_reader_next_packet", "ebiv_reader_seek",
    "ebiv_reader_close", "ebiv_frame_time", "ebiv_frame_index_at",
)


class FileHeader(C.Structure):
    _fields_ = [("version", C.c_uint16), ("tile_cols", C.c_uint16), ("tile_rows", C.c_uint16),
                ("reserved", C.c_uint16), ("width", C.c_int32), ("height", C.c_int32),
                ("fps_num", C.c_uint32), ("fps_den", C.c_uint32), ("chroma_format", C.c_uint8),
                ("bit_depth", C.c_uint8), ("color_range", C.c_uint8), ("color_primaries", C.c_uint8),
                ("flags", C.c_uint32)]


class IndexEntry(C.Structure):
    _fields_ = [("offset", C.c_uint64), ("size", C.c_uint32), ("keyframe", C.c_int32)]


class Schedule(C.Structure):
    _fields_ = [("intra_waves_key", C.c_int32), ("intra_waves_few", C.c_int32),
                ("intra_waves_many", C.c_int32), ("many_tiles", C.c_int32), ("recon_lanes", C.c_int32),
                ("pipeline", C.c_int32), ("intra_waves_key_many", C.c_int32), ("intra_cells", C.c_int32),
                ("key_band_rows", C.c_int32)]


class Frame(C.Structure):
    _fields_ = [("y", C.c_void_p), ("cb", C.c_void_p), ("cr", C.c_void_p),
                ("y_stride", C.c_int32), ("c_stride", C.c_int32),
                ("width", C.c_int32), ("height", C.c_int32),
                ("cwidth", C.c_int32), ("cheight", C.c_int32),
                ("timestamp_ns", C.c_int64), ("on_device", C.c_int32), ("ring_slot", C.c_int32)]


class Completion(C.Structure):
    _fields_ = [("session", C.c_int32), ("status", C.c_int32), ("tag", C.c_uint64), ("frame", Frame),
                ("message", C.c_char * 96), ("surface", C.c_int32), ("reserved", C.c_int32)]


class KernelTimes(C.Structure):
    _fields_ = [("parse_ms", C.c_double), ("inter_ms", C.c_double), ("intra_ms", C.c_double),
                ("rgba_ms", C.c_double), ("parse_n", C.c_uint64), ("inter_n", C.c_uint64),
                ("intra_n", C.c_uint64), ("rgba_n", C.c_uint64), ("inter_mbs", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("symbols", C.c_uint64), ("tiles", C.c_uint64),
                ("tiles_generic", C.c_uint64)]


def bind(L):
    vp, sz = C.c_void_p, C.c_size_t
    L.ebiv_codec_create.restype = C.c_int
    L.ebiv_codec_create.argtypes = [C.c_int, C.c_int, C.POINTER(vp)]
    L.ebiv_codec_decode.restype = C.c_int
    L.ebiv_codec_decode.argtypes = [vp, C.c_char_p, sz, C.c_int64, C.POINTER(Frame)]
    L.ebiv_codec_decode_device.restype = C.c_int
    L.ebiv_codec_decode_device.argtypes = [vp, C.c_char_p, vp, sz, C.c_int64, C.POINTER(Frame)]
    L.ebiv_codec_flush.restype = None
    L.ebiv_codec_flush.argtypes = [vp]
    L.ebiv_codec_destroy.restype = None
    L.ebiv_codec_destroy.argtypes = [vp]
    L.ebiv_codec_last_error.restype = C.c_char_p
    L.ebiv_codec_last_error.argtypes = [vp]
    L.ebiv_codec_stats.restype = C.c_int
    L.ebiv_codec_stats.argtypes = [vp, C.POINTER(Stats)]
    L.ebiv_frame_rgba.restype = C.c_int
    L.ebiv_frame_rgba.argtypes = [vp, C.POINTER(Frame), vp, C.c_int]
    L.ebiv_frame_rgba_noalloc.restype = C.c_int
    L.ebiv_frame_rgba_noalloc.argtypes = [vp, C.POINTER(Frame), vp]
    L.ebiv_device_alloc.restype = C.c_int
    L.ebiv_device_alloc.argtypes = [C.c_int, sz, C.POINTER(vp)]
    L.ebiv_device_free.restype = C.c_int
    L.ebiv_device_free.argtypes = [vp]
    L.ebiv_memcpy_h2d.restype = C.c_int
    L.ebiv_memcpy_h2d.argtypes = [vp, vp, sz]
    L.ebiv_memcpy_d2h.restype = C.c_int
    L.ebiv_memcpy_d2h.argtypes = [vp, vp, sz]
    L.ebiv_device_count.restype = C.c_int
    L.ebiv_device_count.argtypes = []
    L.ebiv_codec_set_timing.restype = C.c_int
    L.ebiv_codec_set_timing.argtypes = [vp, C.c_int]
    L.ebiv_codec_debug_dump.restype = C.c_int
    L.ebiv_codec_debug_dump.argtypes = [vp, C.c_int, vp, sz, vp, sz, vp, sz]
    L.ebiv_codec_kernel_times.restype = C.c_int
    L.ebiv_codec_kernel_times.argtypes = [vp, C.POINTER(KernelTimes)]
    L.ebiv_batch_create.restype = C.c_int
    L.ebiv_batch_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.ebiv_batch_destroy.restype = None
    L.ebiv_batch_destroy.argtypes = [vp]
    L.ebiv_batch_last_error.restype = C.c_char_p
    L.ebiv_batch_last_error.argtypes = [vp]
    L.ebiv_batch_open.restype = C.c_int
    L.ebiv_batch_open.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.ebiv_batch_close.restype = C.c_int
    L.ebiv_batch_close.argtypes = [vp, C.c_int]
    L.ebiv_batch_submit.restype = C.c_int
    L.ebiv_batch_submit.argtypes = [vp, C.c_int, C.c_char_p, vp, sz, C.c_int64, C.c_uint64]
    L.ebiv_batch_flush.restype = C.c_int
    L.ebiv_batch_flush.argtypes = [vp, C.c_int]
    L.ebiv_batch_pending.restype = C.c_int
    L.ebiv_batch_pending.argtypes = [vp, C.c_int]
    L.ebiv_batch_step.restype = C.c_int
    L.ebiv_batch_step.argtypes = [vp, C.POINTER(Completion), C.c_int, C.POINTER(C.c_int)]
    L.ebiv_batch_stats.restype = C.c_int
    L.ebiv_batch_stats.argtypes = [vp, C.c_int, C.POINTER(Stats)]
    L.ebiv_batch_set_timing.restype = C.c_int
    L.ebiv_batch_set_timing.argtypes = [vp, C.c_int]
    L.ebiv_batch_kernel_times.restype = C.c_int
    L.ebiv_batch_kernel_times.argtypes = [vp, C.POINTER(KernelTimes)]
    L.ebiv_batch_set_surfaces.restype = C.c_int
    L.ebiv_batch_set_surfaces.argtypes = [vp, C.c_int, C.POINTER(vp), C.c_int, sz]
    L.ebiv_batch_set_schedule.restype = C.c_int
    L.ebiv_batch_set_schedule.argtypes = [vp, C.POINTER(Schedule)]
    L.ebiv_batch_get_schedule.restype = C.c_int
    L.ebiv_batch_get_schedule.argtypes = [vp, C.POINTER(Schedule)]
    L.ebiv_batch_frame_rgba.restype = C.c_int
    L.ebiv_batch_frame_rgba.argtypes = [vp, C.POINTER(Frame), vp, C.c_int]
    L.ebiv_session_bytes.restype = sz
    L.ebiv_session_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.ebiv_batch_submit_many.restype = C.c_int
    L.ebiv_batch_submit_many.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                         C.POINTER(sz), C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]
    L.ebiv_reader_open.restype = C.c_int
    L.ebiv_reader_open.argtypes = [C.c_char_p, sz, C.POINTER(vp), C.c_char_p, sz]
    L.ebiv_reader_destroy.restype = None
    L.ebiv_reader_destroy.argtypes = [vp]
    L.ebiv_reader_header.restype = C.c_int
    L.ebiv_reader_header.argtypes = [vp, C.POINTER(FileHeader)]
    L.ebiv_reader_frame_count.restype = C.c_int64
    L.ebiv_reader_frame_count.argtypes = [vp]
    L.ebiv_reader_index.restype = C.c_int
    L.ebiv_reader_index.argtypes = [vp, C.POINTER(IndexEntry), sz, C.POINTER(sz)]
    L.ebiv_reader_index_scanned.restype = C.c_int
    L.ebiv_reader_index_scanned.argtypes = [vp]
    L.ebiv_reader_duration.restype = C.c_int64
    L.ebiv_reader_duration.argtypes = [vp]
    L.ebiv_reader_next_packet.restype = C.c_int
    L.ebiv_reader_next_packet.argtypes = [vp, C.POINTER(C.c_void_p), C.POINTER(sz), C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int32)]
    L.ebiv_reader_seek.restype = C.c_int
    L.ebiv_reader_seek.argtypes = [vp, C.c_int64, C.POINTER(C.c_int64)]
    L.ebiv_reader_close.restype = C.c_int
    L.ebiv_reader_close.argtypes = [vp]
    L.ebiv_frame_time.restype = C.c_int64
    L.ebiv_frame_time.argtypes = [C.c_uint32, C.c_uint32, C.c_int64]
    L.ebiv_frame_index_at.restype = C.c_int64
    L.ebiv_frame_index_at.argtypes = [C.c_uint32, C.c_uint32, C.c_int64]
    return L
