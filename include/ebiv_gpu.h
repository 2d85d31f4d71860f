/*
 * ebiv_gpu.h — C ABI of the MI355X (gfx950) EBIV decoder, libebivgpu.so.
 *
 * The entry points are what a cgo shim implementing govid's Codec interface
 * (govid.go:41-44: Decode(Packet) (*Frame, error); Flush()) binds.  Plain
 * pointers and sizes only.  Return values are status codes that map one to one
 * onto ebiv's sentinel errors (ebiv/format.go:110-119).
 *
 * Threading: a codec handle is not thread-safe, like ebiv.Codec
 * (player.go:16-17, source.go:88-95); distinct handles may be used from
 * distinct threads.
 */
#ifndef EBIV_GPU_H
#define EBIV_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  EBIV_ERR_CORRUPT = ErrCorrupt, EBIV_ERR_UNSUPPORTED =
 * ErrUnsupported, EBIV_ERR_NOKEYFRAME = ErrNoKeyframe, EBIV_ERR_DIMENSIONS =
 * ErrDimensions; the rest are runtime failures with no Go counterpart. */
enum {
    EBIV_OK = 0,
    EBIV_ERR_CORRUPT = 1,
    EBIV_ERR_UNSUPPORTED = 2,
    EBIV_ERR_NOKEYFRAME = 3,
    EBIV_ERR_DIMENSIONS = 4,
    EBIV_ERR_BAD_MAGIC = 5, /* ErrBadMagic (container) */
    EBIV_ERR_VERSION = 6,   /* ErrVersion (container) */
    EBIV_EOF = 7,           /* io.EOF from the container reader */
    EBIV_ERR_ARGUMENT = 16, /* null handle, bad pointer or size */
    EBIV_ERR_DEVICE = 17,   /* HIP runtime failure (allocation, launch) */
};

typedef struct ebiv_codec ebiv_codec;

/* A decoded frame: govid.Frame{YCbCr, Timestamp, Width, Height}
 * (govid.go:16-22).  The planes alias the codec's frame ring and stay valid
 * until ring_size further decodes (ebiv/codec.go:15-22).  on_device tells
 * whether the plane pointers are host memory (ebiv_codec_decode) or HBM
 * (ebiv_codec_decode_device). */
typedef struct {
    const uint8_t *y, *cb, *cr;
    int32_t y_stride, c_stride;
    int32_t width, height;   /* luma, as Frame.Width/Height */
    int32_t cwidth, cheight; /* chroma plane size ((w+1)/2, (h+1)/2) */
    int64_t timestamp_ns;    /* Packet.Timestamp, passed through */
    int32_t on_device;
    int32_t ring_slot;
} ebiv_frame;

typedef struct {
    uint64_t frames;       /* successful decodes */
    uint64_t symbols;      /* rANS decode calls, bypass bits included */
    uint64_t tiles;        /* entropy tiles parsed */
    uint64_t tiles_generic; /* of those, parsed by the generic grammar decoder
                              (a frame with no sign table: the hand-scheduled
                              token loop needs one) */
} ebiv_stats;

/* ebiv.NewCodec(WithFrameRing(ring_size)) — ebiv/codec.go:45-64.  ring_size
 * < 2 keeps the default 3.  device selects the HIP device. */
int ebiv_codec_create(int ring_size, int device, ebiv_codec **out);

/* Codec.Decode — ebiv/codec.go:70-111.  data/len is the packet
 * (Packet.Data); it is consumed before the call returns.  On success *out
 * describes host-memory planes (the visible region, strides as the Go ring
 * images: alignUp(w, 64), alignUp(cw, 64)). */
int ebiv_codec_decode(ebiv_codec *c, const uint8_t *data, size_t len, int64_t timestamp_ns, ebiv_frame *out);

/* The same decode for a packet already resident in HBM: host_data is read
 * for the headers and tables, dev_data (same bytes, device memory) by the
 * kernels, which read tile streams as aligned dwords: dev_data must stay
 * readable for 4 bytes past len (buffers from ebiv_device_alloc are).  *out
 * describes HBM planes (padded macroblock-aligned planes; the visible region
 * starts at offset 0). */
int ebiv_codec_decode_device(ebiv_codec *c, const uint8_t *host_data, const uint8_t *dev_data, size_t len,
                             int64_t timestamp_ns, ebiv_frame *out);

/* Codec.Flush — ebiv/codec.go:115-117: resets only the ring position. */
void ebiv_codec_flush(ebiv_codec *c);

void ebiv_codec_destroy(ebiv_codec *c);

/* Message of the last failed call on this handle. */
const char *ebiv_codec_last_error(const ebiv_codec *c);

int ebiv_codec_stats(const ebiv_codec *c, ebiv_stats *out);

/* Per-kernel device time of this codec's decodes, measured with HIP events
 * on the codec's stream while timing is enabled (no reference counterpart:
 * measurement only).  Durations in milliseconds, summed over launches;
 * inter_mbs: macroblocks covered by the timed inter launches. */
typedef struct {
    double parse_ms, inter_ms, intra_ms, rgba_ms;
    uint64_t parse_n, inter_n, intra_n, rgba_n;
    uint64_t inter_mbs;
} ebiv_kernel_times;

int ebiv_codec_set_timing(ebiv_codec *c, int enabled);
int ebiv_codec_kernel_times(const ebiv_codec *c, ebiv_kernel_times *out);

/* Diagnostics (no reference counterpart): copies the entropy kernel's
 * output of the last coded decode to host memory — macroblock records
 * (48 bytes each, frame raster order), the coefficient list, and the per-tile
 * status (16 bytes each).  slot 0 is the frame's payload, 1 a golden override.
 * Returns the number of macroblocks, or a negative status. */
int ebiv_codec_debug_dump(ebiv_codec *c, int slot, void *recs, size_t recs_bytes, uint32_t *coefs,
                          size_t coef_count, void *status, size_t status_bytes);

/* Frame.RGBA for a frame this codec returned — frame.go:8-67, bit-exact.
 * dst receives width*height*4 bytes; dst_on_device selects HBM or host. */
int ebiv_frame_rgba(ebiv_codec *c, const ebiv_frame *f, uint8_t *dst, int dst_on_device);

/* Frame.RGBA of a frame this codec returned into host memory (dst receives
 * width*height*4 bytes), allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (HIP graph
 * capture, tight serving loops).  The conversion is staged in a fixed device
 * buffer of 3840*2160*4 bytes that the library holds on each device it runs on
 * until the process ends; larger frames return EBIV_ERR_ARGUMENT.  Every call
 * on one device shares that buffer: calls that overlap on one device (from
 * several threads or codecs) race. */
int ebiv_frame_rgba_noalloc(ebiv_codec *c, const ebiv_frame *f, uint8_t *dst);

/* ---- Batched decode -----------------------------------------------------------
 *
 * Many streams decoded together on one device.  govid drives one Codec per
 * stream from a frame source (source.go:9-33: next, seek, release; the
 * decode-ahead source of source.go:83-131 keeps a bounded queue of decoded
 * frames); here one engine decodes every stream's queue.  Each session is one
 * ebiv.Codec with the same state machine, errors and ring rules as
 * ebiv_codec_*; a step decodes up to `window` queued packets of every session
 * with one entropy launch over all their tiles and one reconstruction launch
 * per frame position, and returns one completion per packet it consumed, in
 * submission order per session.  Completed frames are HBM planes and stay
 * valid until ring_size further frames of their session have completed.
 * Not thread-safe; distinct engines may run on distinct threads. */
typedef struct ebiv_batch ebiv_batch;

typedef struct {
    int32_t session;
    int32_t status;     /* EBIV_OK or the decode's error (as ebiv_codec_decode) */
    uint64_t tag;       /* the submit call's tag */
    ebiv_frame frame;   /* valid when status == EBIV_OK */
    char message[96];   /* error text when status != EBIV_OK */
    int32_t surface;    /* presentation surface the frame's RGBA was written to, or -1 */
    int32_t reserved;
} ebiv_completion;

int ebiv_batch_create(int device, ebiv_batch **out);
void ebiv_batch_destroy(ebiv_batch *b);
const char *ebiv_batch_last_error(const ebiv_batch *b);

/* Opens a session (a NewCodec(WithFrameRing(ring_size))); window is the most
 * frames of this session one step decodes (1..64).  ring_size is raised to
 * window when smaller. */
int ebiv_batch_open(ebiv_batch *b, int ring_size, int window, int *session);
int ebiv_batch_close(ebiv_batch *b, int session);

/* Queues a packet.  host_data is copied; dev_data, when not null, holds the
 * same bytes in HBM, readable for 4 bytes past len (as ebiv_device_alloc's
 * buffers are), and must stay valid until the packet completes; when null
 * the engine stages the packet itself. */
int ebiv_batch_submit(ebiv_batch *b, int session, const uint8_t *host_data, const uint8_t *dev_data, size_t len,
                      int64_t timestamp_ns, uint64_t tag);
/* n packets of one session in one call, in order (as n ebiv_batch_submit
 * calls; no counterpart in the reference).  dev_data, ts and tags may be
 * null (no HBM copies, timestamps 0, tags 0); nothing is queued when an
 * argument is invalid. */
int ebiv_batch_submit_many(ebiv_batch *b, int session, int n, const uint8_t *const *host_data,
                           const uint8_t *const *dev_data, const size_t *len, const int64_t *ts,
                           const uint64_t *tags);

/* Codec.Flush for a session (codec.go:115-117), dropping its queued packets:
 * the seek path (source.go:45-50 and 226-234 seek the demuxer, then flush).
 * Packets a step in flight has taken were decoded before the flush, in
 * submission order, and complete as usual. */
int ebiv_batch_flush(ebiv_batch *b, int session);

/* Packets of one session (or of all when session is 0) not yet returned:
 * queued, taken by a step in flight, or decoded and held for the next call. */
int ebiv_batch_pending(const ebiv_batch *b, int session);

/* Runs one step; writes at most max_out completions (a step's completions
 * beyond it come first from the next call).  When max_out is less than the
 * queued work, the sessions served first rotate from step to step.  With the
 * pipeline on (ebiv_schedule), a call also takes the next step's packets from
 * the queues and starts their entropy decoding, which the next call
 * finishes: packets submitted after a call join the step after the one in
 * flight.  Returns a runtime status (decode errors are reported per
 * completion). */
int ebiv_batch_step(ebiv_batch *b, ebiv_completion *out, int max_out, int *n_out);

/* Zero-copy presentation.  govid's Ebitengine bridge uploads each frame's
 * Frame.RGBA() to a GPU image (ebitengine/video_image.go:30-52); here the
 * batch step that decodes a session's frames also writes their RGBA
 * (frame.go:8-67, bit-exact) straight into caller-owned device surfaces, in
 * one launch for the whole step.  surfaces: count device pointers (at least
 * the session's window, at most 64), each with room for height rows of pitch
 * bytes; the session's successful
 * frames take them in turn, and a completion's `surface` names the one its
 * frame went to (-1: none, also when pitch < 4 * width).  A surface is
 * rewritten count frames later.  count 0 turns presentation off.  No
 * counterpart in the reference's API (govid cannot preset Frame.RGBA from
 * outside its package, frame.go:76-91). */
int ebiv_batch_set_surfaces(ebiv_batch *b, int session, uint8_t *const *surfaces, int count, size_t pitch);

int ebiv_batch_stats(const ebiv_batch *b, int session, ebiv_stats *out);
int ebiv_batch_set_timing(ebiv_batch *b, int enabled);
int ebiv_batch_kernel_times(const ebiv_batch *b, ebiv_kernel_times *out);
int ebiv_batch_frame_rgba(ebiv_batch *b, const ebiv_frame *f, uint8_t *dst, int dst_on_device);

/* Launch shapes of the batch engine.  No counterpart in the reference (its
 * tile pool, ebiv/pool.go:35-72, has no launch shape); for callers that tune
 * the engine to their mix (govid_amd/tune.py).
 * intra_waves_key and intra_waves_key_many: wavefronts per key-frame tile,
 * intra_waves_few and intra_waves_many: per inter-frame tile, each pair for
 * launches of fewer than / at least many_tiles tiles (1-16).  recon_lanes: HIP streams (1-4) the
 * sessions' reconstruction launches are dealt to after the step's entropy
 * launch, so that one group's launches overlap another's.  pipeline: 1
 * (default) queues the next step's entropy launch on a second stream, where
 * it runs beside this step's reconstruction; 2 runs one step at a time.
 * intra_cells: 1 reconstructs key-frame tiles and tiles dense with intra
 * macroblocks by levels of blocks (tiles that fit: at most 64 x 64 and
 * 1024 macroblocks, and at most as many rows as wavefronts or fewer than
 * twice as many columns), 2 (default) by macroblock wavefronts.
 * key_band_rows: macroblock rows per workgroup of a key-frame tile in a
 * step of fewer than many_tiles key-frame tiles (default 4: a tall tile's
 * rows spread over that many more CUs, each band waiting on the band above
 * through device memory); 128 or more keeps one workgroup per tile.  A
 * key-frame launch of banded tiles runs at most key_band_rows wavefronts
 * per workgroup.  A
 * field <= 0 keeps its value.  Results do not depend on the schedule. */
typedef struct {
    int32_t intra_waves_key, intra_waves_few, intra_waves_many, many_tiles, recon_lanes;
    int32_t pipeline;
    int32_t intra_waves_key_many;
    int32_t intra_cells;
    int32_t key_band_rows;
} ebiv_schedule;
int ebiv_batch_set_schedule(ebiv_batch *b, const ebiv_schedule *s);
int ebiv_batch_get_schedule(const ebiv_batch *b, ebiv_schedule *out);

/* HBM one session holds at this geometry: frame pool and payload buffers
 * (for admission against a device memory budget). */
size_t ebiv_session_bytes(int width, int height, int ring_size, int window);

/* ---- Container reader (host only; needs no GPU) ----
 * ebiv.NewDemuxer / Reader (ebiv/reader.go:32-258) over an in-memory copy of
 * an EBIV file: the 48-byte file header (ebiv/format.go:151-247), the footer
 * index with its CRC-32, and the record scan that rebuilds a missing or
 * damaged index.  Errors: EBIV_ERR_BAD_MAGIC, EBIV_ERR_VERSION,
 * EBIV_ERR_CORRUPT, EBIV_ERR_UNSUPPORTED as Go's sentinels; EBIV_EOF as
 * io.EOF.  Packet data stays valid until ebiv_reader_destroy. */
typedef struct ebiv_reader ebiv_reader;

typedef struct { /* FileHeader (ebiv/format.go:156-169) */
    uint16_t version, tile_cols, tile_rows, reserved;
    int32_t width, height;
    uint32_t fps_num, fps_den;
    uint8_t chroma_format, bit_depth, color_range, color_primaries;
    uint32_t flags;
} ebiv_file_header;

typedef struct { /* IndexEntry (ebiv/format.go:281-285) */
    uint64_t offset; /* of the frame record, prefix included */
    uint32_t size;   /* payload bytes */
    int32_t keyframe;
} ebiv_index_entry;

/* NewDemuxer (reader.go:32-61); msg (may be null) receives the error text. */
int ebiv_reader_open(const uint8_t *data, size_t len, ebiv_reader **out, char *msg, size_t msg_cap);
void ebiv_reader_destroy(ebiv_reader *r);
int ebiv_reader_header(const ebiv_reader *r, ebiv_file_header *out);   /* FileHeader, reader.go:163 */
int64_t ebiv_reader_frame_count(const ebiv_reader *r);                /* FrameCount, reader.go:166 */
/* Index (reader.go:170): copies at most max entries, *n = the count. */
int ebiv_reader_index(const ebiv_reader *r, ebiv_index_entry *out, size_t max, size_t *n);
int ebiv_reader_index_scanned(const ebiv_reader *r); /* 1: index rebuilt by the record scan */
int64_t ebiv_reader_duration(const ebiv_reader *r);  /* Duration, reader.go:183-184 (ns) */
/* NextPacket (reader.go:189-217): EBIV_OK, or EBIV_EOF past the end or
 * after close (sticky). */
int ebiv_reader_next_packet(ebiv_reader *r, const uint8_t **data, size_t *len, int64_t *ts_ns, int32_t *keyframe);
/* Seek (reader.go:222-238): to the last key frame at or before t_ns. */
int ebiv_reader_seek(ebiv_reader *r, int64_t t_ns, int64_t *actual_ns);
int ebiv_reader_close(ebiv_reader *r); /* Close (reader.go:242-245) */
/* frameTime / frameIndexAt (ebiv/format.go:250-275), nanoseconds. */
int64_t ebiv_frame_time(uint32_t fps_num, uint32_t fps_den, int64_t i);
int64_t ebiv_frame_index_at(uint32_t fps_num, uint32_t fps_den, int64_t t_ns);

/* Device memory helpers for callers that keep packets resident in HBM. */
int ebiv_device_alloc(int device, size_t bytes, void **out);
int ebiv_device_free(void *p);
int ebiv_memcpy_h2d(void *dst, const void *src, size_t bytes);
int ebiv_memcpy_d2h(void *dst, const void *src, size_t bytes);

/* Number of HIP devices (0 when none is usable). */
int ebiv_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
