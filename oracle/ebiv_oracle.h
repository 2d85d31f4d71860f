/*
 * ebiv_oracle.h — CPU oracle for the EBIV decode path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of govid's Go
 * `ebiv` decoder (ebiv/codec.go, ebiv/decode_intra.go, ebiv/decode_inter.go,
 * ebiv/rans.go, ebiv/residual.go, ebiv/quant.go, ebiv/transform.go,
 * ebiv/predict.go, ebiv/mv.go, ebiv/motioncomp.go, ebiv/frame.go,
 * ebiv/format.go) and of the root package's YCbCr->RGBA conversion
 * (frame.go:8-67).  It is the checker the GPU decoder is compared against and
 * the CPU baseline of the benchmark.  Nothing in the product path may link or
 * call it; only tests, the smoke check and the benchmark's CPU-baseline leg do.
 *
 * Status codes mirror the Go sentinel errors (ebiv/format.go:110-119).
 */
#ifndef EBIV_ORACLE_H
#define EBIV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_ERR_CORRUPT = 1,     /* ErrCorrupt */
    ORC_ERR_UNSUPPORTED = 2, /* ErrUnsupported */
    ORC_ERR_NOKEYFRAME = 3,  /* ErrNoKeyframe */
    ORC_ERR_DIMENSIONS = 4,  /* ErrDimensions */
};

typedef struct orc_codec orc_codec;

typedef struct {
    const uint8_t *y, *cb, *cr; /* ring image planes (codec-owned) */
    int y_stride, c_stride;
    int width, height;          /* luma size */
    int cwidth, cheight;        /* chroma size */
} orc_frame;

/* NewCodec(WithFrameRing(ring)) — ebiv/codec.go:45-64. ring < 2 keeps 3. */
orc_codec *orc_codec_new(int ring);
void orc_codec_free(orc_codec *c);
/* Codec.Decode — ebiv/codec.go:70-111. Returns an ORC_* status. */
int orc_codec_decode(orc_codec *c, const uint8_t *data, size_t len, orc_frame *out);
/* Codec.Flush — ebiv/codec.go:115-117. */
void orc_codec_flush(orc_codec *c);
/* Message of the last failed decode. */
const char *orc_codec_last_error(const orc_codec *c);
/* Where the last decode failed: 0 not in tiles (or no failure), 1 in the
 * frame's own payload's tiles, 2 in its golden override's tiles. */
int orc_codec_last_tile_failure(const orc_codec *c);
/* Worker threads for tile decode (ebiv/pool.go: NumCPU workers; 1 = inline). */
void orc_codec_set_threads(orc_codec *c, int threads);
/* Number of rANS decode calls (bypass bits included) made by this codec. */
uint64_t orc_codec_symbols(const orc_codec *c);
/* The golden reference buffer (padded planes), for state checks in tests. */
int orc_codec_golden(const orc_codec *c, const uint8_t **y, const uint8_t **cb,
                     const uint8_t **cr, int *ypw, int *yph);

/* Per-macroblock trace of the last coded decode (the frame's own payload, not
 * a golden override): what the grammar read, for comparing the GPU's parse.
 * coef entries: bits 0-7 position, 8-12 block (0-15 luma in decode order,
 * 16 Cb, 17 Cr), 13-31 signed level. */
typedef struct {
    int32_t kind;              /* 0 skip, 1 inter, 2 intra, -1 not decoded */
    int32_t tx, cmode, refsel, part, cbp;
    int32_t lmode[16];
    int16_t mv[4][2];
    int32_t ncoef;
    uint32_t coef[384];
} orc_mbtrace;

/* Enables tracing (allocates one record per macroblock on the next key
 * frame); orc_codec_trace returns the records and their count. */
void orc_codec_enable_trace(orc_codec *c, int on);
const orc_mbtrace *orc_codec_trace(const orc_codec *c, int *count);

/* convertYCbCr420ToRGBA — frame.go:8-67 (Rect.Min = 0). */
void orc_ycbcr_to_rgba(const uint8_t *y, int y_stride, const uint8_t *cb,
                       const uint8_t *cr, int c_stride, int w, int h, uint8_t *dst);

/* ---- unit-level entry points for known-answer tests ---- */
/* inverseDCT — ebiv/transform.go:129-194. n in {4,8,16}. */
int orc_inverse_dct(const int32_t *src, int32_t *dst, int n);
/* dctMatrix entry C[k][i] — ebiv/transform.go:27-40. */
int32_t orc_dct_coef(int n, int k, int i);
/* quantStep[qp] — ebiv/quant.go:19-29. */
int32_t orc_quant_step(int qp);
/* scanOrders[n][i] — ebiv/scan.go:17-35. */
int orc_zigzag(int n, int i);
/* predictIntra + reconstruct-free prediction — ebiv/predict.go:63-129.
 * plane is w x h with the given stride; pred receives n*n int32. */
int orc_predict_intra(const uint8_t *plane, int stride, int w, int h, int x0,
                      int y0, int n, int mode, int av_above, int av_left,
                      int av_ar, int32_t *pred);
/* aboveRightAvailable — ebiv/predict.go:50-58. */
int orc_above_right_available(int x0, int y0, int n, int tile_top, int tile_right,
                              int mb_top, int mb_right);
/* mcLuma / mcChroma (int32 output), ebiv/motioncomp.go:112-123, 214-256. */
void orc_mc_luma(const uint8_t *plane, int stride, int w, int h, int src_x, int src_y,
                 int fx, int fy, int pw, int ph, int32_t *out, int out_stride);
void orc_mc_chroma(const uint8_t *plane, int stride, int w, int h, int src_x, int src_y,
                   int fx, int fy, int pw, int ph, int32_t *out, int out_stride);
/* predictMV over an mv grid of (x,y) int16 pairs — ebiv/mv.go:21-52. */
void orc_predict_mv(const int16_t *mv, int stride, int mbx, int mby, int x0, int x1,
                    int y0, int y1, int16_t *out);
/* Decode a sequence of symbols from one rANS tile stream.  tables_blob is a
 * serialized key-frame table set (ebiv/rans.go:396-471); ctxs[i] is the
 * context of symbol i.  Returns ORC_OK or the decode error. */
int orc_rans_decode_symbols(const uint8_t *tables_blob, size_t tables_len,
                            const uint8_t *stream, size_t stream_len,
                            const int *ctxs, int count, int *out_syms);
/* decodeBlock over a tile stream: decodes `count` consecutive blocks of size
 * n (plane/txIdx fixed) into levels (count*n*n int32, row-major per block). */
int orc_decode_blocks(const uint8_t *tables_blob, size_t tables_len,
                      const uint8_t *stream, size_t stream_len, int plane,
                      int tx_idx, int n, int count, int32_t *levels);

#ifdef __cplusplus
}
#endif
#endif
