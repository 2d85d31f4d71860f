/*
 * ebiv_oracle.c — CPU oracle for the EBIV decode path (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of govid's Go EBIV decoder, used only by tests, the
 * smoke check and the benchmark's CPU-baseline leg; the product decoder never
 * links it.  Every function cites the Go file:line it restates.  Where Go would
 * panic on a hostile stream (slice bounds after an int() wrap, a slot2sym
 * index past 4096), the oracle returns ORC_ERR_CORRUPT instead.
 */
#include "ebiv_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- coding model: ebiv/model.go:20-213 ---------------------------------- */
enum { MBS = 16, CMBS = 8, DC_PRED = 128 };
enum { M_DC, M_V, M_H, M_TM, M_D45, M_D27, M_D23, M_D135, M_D117, M_D113, N_INTRA_MODES };
enum { MB_SKIP, MB_INTER, MB_INTRA };
enum { REF_LAST, REF_GOLDEN };
enum { CBP_LUMA = 1, CBP_CB = 2, CBP_CR = 4 };
enum { T_ZERO, T_ONE, T_TWO, T_THREE, T_FOUR, T_ESC, T_EOB };
enum {
    CTX_TXSIZE, CTX_LUMAMODE, CTX_CHROMAMODE, CTX_MBMODE, CTX_REF, CTX_PART,
    CTX_CBP, CTX_SIGN, CTX_ESC, CTX_MVX, CTX_MVY, CTX_BYPASS, CTX_TOKEN0
};
#define N_BANDS 5
#define N_PREV 3
#define LUMA_TOK_CTX (3 * N_BANDS * N_PREV)
#define N_CTX (CTX_TOKEN0 + LUMA_TOK_CTX + N_BANDS * N_PREV) /* 72 */
#define RANS_M 4096u
#define RANS_LOWER (1u << 23)
#define MAX_QP 63
#define SDH_MIN_SPAN 4
#define MV_CLAMP 4096
#define MAX_DECODE_PIXELS (8192 * 4320)

static int alphabet(int c) /* alphabetSizes, model.go:195-213 */
{
    static const int a[CTX_TOKEN0] = {3, 10, 10, 3, 2, 4, 8, 2, 16, 16, 16, 2};
    return c < CTX_TOKEN0 ? a[c] : 7;
}

/* partRects, model.go:91-96: {x, y, w, h} in luma pixels. */
static const int part_count[4] = {1, 2, 2, 4};
static const int part_rect[4][4][4] = {
    {{0, 0, 16, 16}},
    {{0, 0, 16, 8}, {0, 8, 16, 8}},
    {{0, 0, 8, 16}, {8, 0, 8, 16}},
    {{0, 0, 8, 8}, {8, 0, 8, 8}, {0, 8, 8, 8}, {8, 8, 8, 8}},
};

/* tokenCtx, model.go:156-161. */
static int token_ctx(int plane, int tx, int band, int prev)
{
    if (plane == 0) return CTX_TOKEN0 + (tx * N_BANDS + band) * N_PREV + prev;
    return CTX_TOKEN0 + LUMA_TOK_CTX + band * N_PREV + prev;
}
/* bandOf, model.go:165-178. */
static int band_of(int i) { return i == 0 ? 0 : i <= 2 ? 1 : i <= 8 ? 2 : i <= 20 ? 3 : 4; }
/* classForMag, model.go:182-191. */
static int class_for_mag(int m) { return m <= 0 ? 0 : m == 1 ? 1 : 2; }
static inline uint8_t clamp_byte(int32_t v) { return v < 0 ? 0 : v > 255 ? 255 : (uint8_t)v; }

/* ---- static tables: scan.go, transform.go, quant.go ---------------------- */
static int zz[3][256];       /* buildZigzag, scan.go:17-35 */
static int32_t dct[3][256];  /* buildDCTMatrix, transform.go:27-40 */
static int32_t qstep[MAX_QP + 1]; /* quantStep, quant.go:19-29 */
static pthread_once_t init_once = PTHREAD_ONCE_INIT;

static int size_idx(int n) { return n == 4 ? 0 : n == 8 ? 1 : 2; }

static void build_static(void)
{
    for (int s = 0; s < 3; s++) {
        int n = 4 << s, k = 0;
        for (int d = 0; d < 2 * n - 1; d++) {
            if (d % 2 == 0) {
                for (int y = d < n - 1 ? d : n - 1; y >= 0 && d - y < n; y--) zz[s][k++] = y * n + (d - y);
            } else {
                for (int x = d < n - 1 ? d : n - 1; x >= 0 && d - x < n; x--) zz[s][k++] = (d - x) * n + x;
            }
        }
        for (int kk = 0; kk < n; kk++) {
            double alpha = kk == 0 ? sqrt(1.0 / n) : sqrt(2.0 / n);
            for (int i = 0; i < n; i++) {
                double c = alpha * cos((2.0 * i + 1.0) * kk * M_PI / (2.0 * n));
                dct[s][kk * n + i] = (int32_t)round(c * 8192.0);
            }
        }
    }
    for (int qp = 0; qp <= MAX_QP; qp++) {
        double st = round(pow(2.0, qp / 6.0));
        qstep[qp] = (int32_t)(st < 1 ? 1 : st);
    }
}
static void ensure_static(void) { pthread_once(&init_once, build_static); }

/* ---- Go encoding/binary varints ------------------------------------------ */
/* binary.Uvarint: n > 0 bytes read, 0 truncated, < 0 overflow. */
static int uvarint(const uint8_t *b, size_t len, uint64_t *out)
{
    uint64_t x = 0;
    unsigned s = 0;
    for (size_t i = 0; i < len; i++) {
        if (i == 10) return -(int)(i + 1);
        uint8_t c = b[i];
        if (c < 0x80) {
            if (i == 9 && c > 1) return -(int)(i + 1);
            *out = x | (uint64_t)c << s;
            return (int)i + 1;
        }
        x |= (uint64_t)(c & 0x7f) << s;
        s += 7;
    }
    *out = 0;
    return 0;
}
static int varint(const uint8_t *b, size_t len, int64_t *out)
{
    uint64_t ux;
    int n = uvarint(b, len, &ux);
    int64_t x = (int64_t)(ux >> 1);
    if (ux & 1) x = ~x;
    *out = x;
    return n;
}

/* ---- rANS tables: rans.go:27-148, 396-481 --------------------------------- */
typedef struct {
    uint32_t start[16], freq[16];
    int used;
    uint8_t slot2sym[RANS_M];
} otable;

typedef struct { /* a prevFreqs slice: backing kept across key-frame wipes */
    uint32_t v[16];
    int len;
    int alloc;
} oprev;

/* buildTableInto (rans.go:120-148).  Returns 0, or -1 where Go would index
 * slot2sym past its end (a wrapped uint32 sum with an entry above ransM). */
static int build_table(otable *t, const uint32_t *f, int n)
{
    uint32_t cum = 0;
    t->used = 0;
    for (int s = 0; s < n; s++) {
        t->start[s] = cum;
        t->freq[s] = f[s];
        cum += f[s];
    }
    if (cum != RANS_M) return 0;
    for (int s = 0; s < n; s++)
        if (f[s] > RANS_M) return -1;
    t->used = 1;
    for (int s = 0; s < n; s++)
        for (uint32_t k = 0; k < f[s]; k++) t->slot2sym[t->start[s] + k] = (uint8_t)s;
    return 0;
}

/* parseTablesInto (rans.go:396-471).  Returns bytes consumed or -1 (corrupt). */
static long parse_tables(otable *tables, oprev *prev, const uint8_t *b, size_t len,
                         int key, const char **msg)
{
    if (key)
        for (int c = 0; c < N_CTX; c++) prev[c].len = 0;
    size_t pos = 0;
    for (int c = 0; c < N_CTX; c++) {
        int n = alphabet(c);
        if (pos >= len) { *msg = "truncated table modes"; return -1; }
        uint8_t mode = b[pos++];
        if (c == CTX_BYPASS && mode != 0) { *msg = "bypass context must not ship a table"; return -1; }
        int have_ref = prev[c].len == n;
        switch (mode) {
        case 0: /* tblUnused */
            tables[c].used = 0;
            break;
        case 1: /* tblSame */
            if (key || !have_ref) { *msg = "context references a missing previous table"; return -1; }
            if (build_table(&tables[c], prev[c].v, n) < 0 || !tables[c].used) {
                *msg = "context frequencies do not sum to 4096"; return -1;
            }
            break;
        case 2: { /* tblFull; growFreqs keeps the old backing */
            if (!prev[c].alloc) { memset(prev[c].v, 0, sizeof prev[c].v); prev[c].alloc = 1; }
            prev[c].len = n;
            for (int s = 0; s < n; s++) {
                uint64_t v;
                int k = uvarint(b + pos, len - pos, &v);
                if (k <= 0) { *msg = "truncated frequency table body"; return -1; }
                pos += (size_t)k;
                prev[c].v[s] = (uint32_t)v;
            }
            if (build_table(&tables[c], prev[c].v, n) < 0 || !tables[c].used) {
                *msg = "context frequencies do not sum to 4096"; return -1;
            }
            break;
        }
        case 3: /* tblDelta */
            if (key || !have_ref) { *msg = "context deltas a missing previous table"; return -1; }
            for (int s = 0; s < n; s++) {
                int64_t v;
                int k = varint(b + pos, len - pos, &v);
                if (k <= 0) { *msg = "truncated frequency table delta"; return -1; }
                pos += (size_t)k;
                int64_t nv = (int64_t)prev[c].v[s] + v;
                if (nv < 0 || nv > (int64_t)RANS_M) { *msg = "context delta out of range"; return -1; }
                prev[c].v[s] = (uint32_t)nv;
            }
            if (build_table(&tables[c], prev[c].v, n) < 0 || !tables[c].used) {
                *msg = "context frequencies do not sum to 4096"; return -1;
            }
            break;
        default:
            *msg = "unknown table mode";
            return -1;
        }
    }
    static const uint32_t bypass[2] = {RANS_M / 2, RANS_M / 2}; /* installBypass, rans.go:317-322 */
    build_table(&tables[CTX_BYPASS], bypass, 2);
    return (long)pos;
}

/* ---- rANS decoder: rans.go:222-310 ---------------------------------------- */
typedef struct {
    const uint8_t *buf;
    size_t pos;
    const otable *tables;
    uint32_t st[4];
    uint32_t counter;
    int err;
    const char *msg;
    uint64_t nsym;
} rdec;

static void rfail(rdec *d, const char *m)
{
    if (!d->err) { d->err = ORC_ERR_CORRUPT; d->msg = m; }
}

static int rdec_reset(rdec *d, const uint8_t *buf, size_t len, const otable *tables)
{
    if (len < 16) return -1;
    d->buf = buf;
    d->pos = len;
    d->tables = tables;
    d->counter = 0;
    d->err = 0;
    d->msg = NULL;
    for (int i = 3; i >= 0; i--) { /* pop32: big-endian from the top */
        uint32_t a = buf[--d->pos], b = buf[--d->pos], c = buf[--d->pos], e = buf[--d->pos];
        d->st[i] = a << 24 | b << 16 | c << 8 | e;
    }
    return 0;
}

static int rdec_decode(rdec *d, int ctx)
{
    if (d->err) return 0;
    const otable *t = &d->tables[ctx];
    if (!t->used) { rfail(d, "symbol decoded from unused context"); return 0; }
    int i = (int)(d->counter & 3);
    d->counter++;
    uint32_t x = d->st[i];
    uint32_t slot = x & (RANS_M - 1);
    int sym = t->slot2sym[slot];
    x = t->freq[sym] * (x >> 12) + slot - t->start[sym];
    while (x < RANS_LOWER) {
        if (d->pos == 0) { rfail(d, "rANS stream underrun"); return 0; }
        x = (x << 8) | d->buf[--d->pos];
    }
    d->st[i] = x;
    d->nsym++;
    return sym;
}

/* getClassValue / getMVComponent / getEscapeValue, mv.go:95-128. */
static unsigned class_value(rdec *d, int ctx)
{
    int cls = rdec_decode(d, ctx);
    if (cls < 0 || cls >= 16) { rfail(d, "class out of range"); return 0; }
    if (cls == 0) return 0;
    if (cls == 1) return 1;
    unsigned v = 0;
    for (int i = 0; i < cls - 1; i++) v = v << 1 | (unsigned)rdec_decode(d, CTX_BYPASS);
    return 1u << (cls - 1) | v;
}
static int zigzag_dec(unsigned u) { return (int)(u >> 1) ^ -(int)(u & 1); } /* motioncomp.go:410-412 */

/* decodeBlock, residual.go:79-126. levels must be zeroed. */
static void decode_block(rdec *d, int plane, int tx, int32_t *levels, int n)
{
    const int *scan = zz[size_idx(n)];
    int total = n * n, first = -1, last = -1, prev = 0;
    int32_t sum_abs = 0;
    for (int i = 0; i < total; i++) {
        int t = rdec_decode(d, token_ctx(plane, tx, band_of(i), prev));
        if (t == T_EOB) break;
        if (t == T_ZERO) { levels[scan[i]] = 0; prev = 0; continue; }
        int mag;
        if (t == T_ESC) { mag = 5 + (int)class_value(d, CTX_ESC); prev = 2; }
        else { mag = t; prev = class_for_mag(mag); }
        sum_abs += mag;
        if (first < 0) first = i;
        else if (rdec_decode(d, CTX_SIGN) == 1) mag = -mag;
        levels[scan[i]] = mag;
        last = i;
    }
    if (first < 0) return;
    int neg = (last - first >= SDH_MIN_SPAN) ? (sum_abs & 1) : rdec_decode(d, CTX_SIGN) == 1;
    if (neg) levels[scan[first]] = -levels[scan[first]];
}

/* inverseDCT, transform.go:129-194 (sparse bounding box, DC-only fast path). */
static void inverse_dct(const int32_t *src, int32_t *dst, int n)
{
    const int32_t *c = dct[size_idx(n)];
    int max_r = -1, max_c = -1;
    for (int r = 0; r < n; r++)
        for (int cc = 0; cc < n; cc++)
            if (src[r * n + cc]) { if (r > max_r) max_r = r; if (cc > max_c) max_c = cc; }
    if (max_r < 0) { memset(dst, 0, sizeof(int32_t) * n * n); return; }
    const int64_t rnd = 4096;
    if (max_r == 0 && max_c == 0) {
        int64_t d0 = c[0];
        int64_t col = (d0 * src[0] + rnd) >> 13;
        int32_t v = (int32_t)((d0 * col + rnd) >> 13);
        for (int i = 0; i < n * n; i++) dst[i] = v;
        return;
    }
    int32_t tmp[256];
    for (int i = 0; i < n; i++)
        for (int l = 0; l <= max_c; l++) {
            int64_t acc = 0;
            for (int k = 0; k <= max_r; k++) acc += (int64_t)c[k * n + i] * src[k * n + l];
            tmp[i * n + l] = (int32_t)((acc + rnd) >> 13);
        }
    for (int i = 0; i < n; i++)
        for (int m = 0; m < n; m++) {
            int64_t acc = 0;
            for (int l = 0; l <= max_c; l++) acc += (int64_t)c[l * n + m] * tmp[i * n + l];
            dst[i * n + m] = (int32_t)((acc + rnd) >> 13);
        }
}

/* ---- planes & intra prediction: predict.go -------------------------------- */
typedef struct { uint8_t *data; int stride, w, h; } plane;
typedef struct { plane y, cb, cr; int mbc, mbr; } fbuf;

typedef struct { int angle, inv, vertical; } angmode;
static const angmode ang_modes[N_INTRA_MODES] = { /* angModes, predict.go:36-43 */
    [M_D45] = {32, 0, 1},   [M_D27] = {13, 0, 1},   [M_D23] = {-13, -630, 1},
    [M_D135] = {32, 0, 0},  [M_D117] = {13, 0, 0},  [M_D113] = {-13, -630, 0},
};

/* aboveRightAvailable, predict.go:50-58. */
static int above_right_avail(int x0, int y0, int n, int tile_top, int tile_right, int mb_top, int mb_right)
{
    if (y0 <= tile_top || x0 + 2 * n > tile_right) return 0;
    if (y0 == mb_top) return 1;
    return x0 + 2 * n <= mb_right;
}

/* substituteRefs, predict.go:134-156. */
static void substitute(int32_t *ref, const int *avail, int len)
{
    int first = -1;
    for (int i = 0; i < len; i++) if (avail[i]) { first = i; break; }
    if (first < 0) { for (int i = 0; i < len; i++) ref[i] = DC_PRED; return; }
    for (int i = 0; i < first; i++) ref[i] = ref[first];
    for (int i = first + 1; i < len; i++) if (!avail[i]) ref[i] = ref[i - 1];
}

/* angularPredict, predict.go:162-203. */
static void angular(const int32_t *ref, int corner, int n, angmode m, int32_t *pred)
{
    int32_t ext[3 * MBS + 2] = {0};
    int off = n;
    for (int k = 0; k <= 2 * n; k++) ext[off + k] = m.vertical ? ref[corner + k] : ref[corner - k];
    if (m.angle < 0) {
        int low = (n * m.angle) >> 5;
        for (int x = -1; x >= low; x--) {
            int idx = (x * m.inv + 128) >> 8;
            ext[off + x] = m.vertical ? ref[corner - idx] : ref[corner + idx];
        }
    }
    for (int a = 0; a < n; a++) {
        int pos = (a + 1) * m.angle, idx = pos >> 5, fact = pos & 31;
        for (int b = 0; b < n; b++) {
            int32_t s0 = ext[off + b + idx + 1], s1 = ext[off + b + idx + 2];
            int32_t v = ((32 - fact) * s0 + fact * s1 + 16) >> 5;
            if (m.vertical) pred[a * n + b] = v; else pred[b * n + a] = v;
        }
    }
}

/* predictIntra, predict.go:63-129. */
static void predict_intra(const plane *p, int x0, int y0, int n, int mode, int av_above,
                          int av_left, int av_ar, int32_t *pred)
{
    int32_t ref[4 * MBS + 1] = {0};
    int avail[4 * MBS + 1] = {0};
    int corner = 2 * n;
    for (int i = 0; i < 2 * n; i++) {
        int row = 2 * n - 1 - i;
        if (av_left && row < n) { ref[i] = p->data[(y0 + row) * p->stride + x0 - 1]; avail[i] = 1; }
    }
    if (av_above && av_left) { ref[corner] = p->data[(y0 - 1) * p->stride + x0 - 1]; avail[corner] = 1; }
    for (int j = 0; j < 2 * n; j++)
        if (av_above && (j < n || av_ar)) {
            ref[corner + 1 + j] = p->data[(y0 - 1) * p->stride + x0 + j];
            avail[corner + 1 + j] = 1;
        }
    substitute(ref, avail, 4 * n + 1);
#define TOP(x) ref[corner + 1 + (x)]
#define LEFT(y) ref[corner - 1 - (y)]
    switch (mode) {
    case M_V:
        for (int r = 0; r < n; r++) for (int c = 0; c < n; c++) pred[r * n + c] = TOP(c);
        break;
    case M_H:
        for (int r = 0; r < n; r++) for (int c = 0; c < n; c++) pred[r * n + c] = LEFT(r);
        break;
    case M_TM:
        for (int r = 0; r < n; r++)
            for (int c = 0; c < n; c++) pred[r * n + c] = clamp_byte(TOP(c) + LEFT(r) - ref[corner]);
        break;
    case M_DC: {
        int32_t sum = 0;
        for (int i = 0; i < n; i++) sum += TOP(i) + LEFT(i);
        int32_t dc = (sum + n) / (2 * n);
        for (int i = 0; i < n * n; i++) pred[i] = dc;
        break;
    }
    default:
        angular(ref, corner, n, ang_modes[mode], pred);
    }
#undef TOP
#undef LEFT
}

/* reconstruct, predict.go:209-219. */
static void reconstruct(plane *p, int x0, int y0, int n, const int32_t *pred, const int32_t *res)
{
    for (int r = 0; r < n; r++) {
        uint8_t *dst = p->data + (y0 + r) * p->stride + x0;
        for (int c = 0; c < n; c++) dst[c] = clamp_byte(pred[r * n + c] + res[r * n + c]);
    }
}

/* ---- motion compensation: motioncomp.go ----------------------------------- */
static const int32_t luma_taps[4][4] = {{0, 64, 0, 0}, {-6, 55, 18, -3}, {-8, 40, 40, -8}, {-3, 18, 55, -6}};
static inline int clampc(int v, int hi) { return v < 0 ? 0 : v >= hi ? hi - 1 : v; }
#define PX(p, x, y) ((int32_t)(p)->data[clampc((y), (p)->h) * (p)->stride + clampc((x), (p)->w)])

/* lumaSubpel, motioncomp.go:49-78. */
static int32_t luma_subpel(const plane *r, int px, int py, int fx, int fy)
{
    if (fx == 0 && fy == 0) return PX(r, px, py);
    if (fy == 0) {
        const int32_t *t = luma_taps[fx];
        int32_t v = t[0] * PX(r, px - 1, py) + t[1] * PX(r, px, py) + t[2] * PX(r, px + 1, py) + t[3] * PX(r, px + 2, py);
        return clamp_byte((v + 32) >> 6);
    }
    if (fx == 0) {
        const int32_t *t = luma_taps[fy];
        int32_t v = t[0] * PX(r, px, py - 1) + t[1] * PX(r, px, py) + t[2] * PX(r, px, py + 1) + t[3] * PX(r, px, py + 2);
        return clamp_byte((v + 32) >> 6);
    }
    const int32_t *tx = luma_taps[fx], *ty = luma_taps[fy];
    int32_t h[4];
    for (int k = 0; k < 4; k++) {
        int yy = py - 1 + k;
        h[k] = tx[0] * PX(r, px - 1, yy) + tx[1] * PX(r, px, yy) + tx[2] * PX(r, px + 1, yy) + tx[3] * PX(r, px + 2, yy);
    }
    int32_t v = ty[0] * h[0] + ty[1] * h[1] + ty[2] * h[2] + ty[3] * h[3];
    return clamp_byte((v + (1 << 11)) >> 12);
}

/* chromaSubpel, motioncomp.go:82-97. */
static int32_t chroma_subpel(const plane *r, int px, int py, int fx, int fy)
{
    if (fx == 0 && fy == 0) return PX(r, px, py);
    int32_t a = PX(r, px, py), b = PX(r, px + 1, py), c = PX(r, px, py + 1), d = PX(r, px + 1, py + 1);
    int32_t v = (8 - fx) * (8 - fy) * a + fx * (8 - fy) * b + (8 - fx) * fy * c + fx * fy * d;
    return (v + 32) >> 6;
}

/* interiorLuma, motioncomp.go:125-130. */
static int interior_luma(const plane *r, int sx, int sy, int fx, int fy, int pw, int ph)
{
    if (fx == 0 && fy == 0) return sx >= 0 && sy >= 0 && sx + pw <= r->w && sy + ph <= r->h;
    return sx - 1 >= 0 && sy - 1 >= 0 && sx + pw + 2 <= r->w && sy + ph + 2 <= r->h;
}

/* Interior luma sample (mcLumaInterior / mcLumaByte interior, motioncomp.go:132-185, 276-323). */
static inline int32_t luma_interior_px(const plane *r, int x, int y, int fx, int fy)
{
    const uint8_t *d = r->data;
    int s = r->stride, b = y * s + x;
    if (fx == 0 && fy == 0) return d[b];
    if (fy == 0) {
        const int32_t *t = luma_taps[fx];
        return clamp_byte((t[0] * d[b - 1] + t[1] * d[b] + t[2] * d[b + 1] + t[3] * d[b + 2] + 32) >> 6);
    }
    if (fx == 0) {
        const int32_t *t = luma_taps[fy];
        return clamp_byte((t[0] * d[b - s] + t[1] * d[b] + t[2] * d[b + s] + t[3] * d[b + 2 * s] + 32) >> 6);
    }
    const int32_t *tx = luma_taps[fx], *ty = luma_taps[fy];
    int32_t h[4];
    for (int k = 0; k < 4; k++) {
        int bb = b + (k - 1) * s;
        h[k] = tx[0] * d[bb - 1] + tx[1] * d[bb] + tx[2] * d[bb + 1] + tx[3] * d[bb + 2];
    }
    return clamp_byte((ty[0] * h[0] + ty[1] * h[1] + ty[2] * h[2] + ty[3] * h[3] + (1 << 11)) >> 12);
}

/* mcLuma, motioncomp.go:112-123 (int32 output). */
static void mc_luma(const plane *r, int sx, int sy, int fx, int fy, int pw, int ph, int32_t *out,
                    int dx, int dy, int ds)
{
    int in = interior_luma(r, sx, sy, fx, fy, pw, ph);
    for (int y = 0; y < ph; y++)
        for (int x = 0; x < pw; x++)
            out[(dy + y) * ds + dx + x] = in ? luma_interior_px(r, sx + x, sy + y, fx, fy)
                                             : luma_subpel(r, sx + x, sy + y, fx, fy);
}

/* mcLumaByte, motioncomp.go:265-324. */
static void mc_luma_byte(const plane *r, int sx, int sy, int fx, int fy, int pw, int ph, plane *dst,
                         int dx, int dy)
{
    int in = interior_luma(r, sx, sy, fx, fy, pw, ph);
    for (int y = 0; y < ph; y++) {
        uint8_t *o = dst->data + (dy + y) * dst->stride + dx;
        for (int x = 0; x < pw; x++)
            o[x] = clamp_byte(in ? luma_interior_px(r, sx + x, sy + y, fx, fy)
                                 : luma_subpel(r, sx + x, sy + y, fx, fy));
    }
}

static int interior_chroma(const plane *r, int sx, int sy, int fx, int fy, int pw, int ph)
{
    if (fx == 0 && fy == 0) return sx >= 0 && sy >= 0 && sx + pw <= r->w && sy + ph <= r->h;
    return sx >= 0 && sy >= 0 && sx + pw + 1 <= r->w && sy + ph + 1 <= r->h;
}
static inline int32_t chroma_interior_px(const plane *r, int x, int y, int fx, int fy)
{
    const uint8_t *d = r->data;
    int s = r->stride, b = y * s + x;
    if (fx == 0 && fy == 0) return d[b];
    int32_t w00 = (8 - fx) * (8 - fy), w01 = fx * (8 - fy), w10 = (8 - fx) * fy, w11 = fx * fy;
    return (w00 * d[b] + w01 * d[b + 1] + w10 * d[b + s] + w11 * d[b + s + 1] + 32) >> 6;
}

/* mcChroma, motioncomp.go:214-256. */
static void mc_chroma(const plane *r, int sx, int sy, int fx, int fy, int pw, int ph, int32_t *out,
                      int dx, int dy, int ds)
{
    int in = interior_chroma(r, sx, sy, fx, fy, pw, ph);
    for (int y = 0; y < ph; y++)
        for (int x = 0; x < pw; x++)
            out[(dy + y) * ds + dx + x] = in ? chroma_interior_px(r, sx + x, sy + y, fx, fy)
                                             : chroma_subpel(r, sx + x, sy + y, fx, fy);
}

/* mcChromaByte, motioncomp.go:329-368. */
static void mc_chroma_byte(const plane *r, int sx, int sy, int fx, int fy, int pw, int ph, plane *dst,
                           int dx, int dy)
{
    int in = interior_chroma(r, sx, sy, fx, fy, pw, ph);
    for (int y = 0; y < ph; y++) {
        uint8_t *o = dst->data + (dy + y) * dst->stride + dx;
        for (int x = 0; x < pw; x++)
            o[x] = clamp_byte(in ? chroma_interior_px(r, sx + x, sy + y, fx, fy)
                                 : chroma_subpel(r, sx + x, sy + y, fx, fy));
    }
}

/* clampMV, motioncomp.go:394-402. */
static int16_t clamp_mv(int v) { return (int16_t)(v < -MV_CLAMP ? -MV_CLAMP : v > MV_CLAMP ? MV_CLAMP : v); }

/* median3 / predictMV, mv.go:21-65.  mv holds (x, y) int16 pairs. */
static int16_t median3(int16_t a, int16_t b, int16_t c)
{
    if (a > b) { int16_t t = a; a = b; b = t; }
    if (b > c) b = c;
    if (a > b) b = a;
    return b;
}
typedef struct { int x0, x1, y0, y1; } tbounds; /* tileBounds in MBs */

static void predict_mv(const int16_t *mv, int stride, int mbx, int mby, tbounds b, int16_t *out)
{
    int16_t a[2] = {0, 0}, bv[2] = {0, 0}, c[2] = {0, 0};
    int aok = mbx > b.x0, bok = mby > b.y0, cok = 0;
    if (aok) { a[0] = mv[2 * (mby * stride + mbx - 1)]; a[1] = mv[2 * (mby * stride + mbx - 1) + 1]; }
    if (bok) {
        bv[0] = mv[2 * ((mby - 1) * stride + mbx)];
        bv[1] = mv[2 * ((mby - 1) * stride + mbx) + 1];
        if (mbx + 1 < b.x1) {
            c[0] = mv[2 * ((mby - 1) * stride + mbx + 1)]; c[1] = mv[2 * ((mby - 1) * stride + mbx + 1) + 1]; cok = 1;
        } else if (aok) {
            c[0] = mv[2 * ((mby - 1) * stride + mbx - 1)]; c[1] = mv[2 * ((mby - 1) * stride + mbx - 1) + 1]; cok = 1;
        }
    }
    if (!aok && !bok && !cok) { out[0] = out[1] = 0; return; }
    if (aok && !bok && !cok) { out[0] = a[0]; out[1] = a[1]; return; }
    if (!aok && bok && !cok) { out[0] = bv[0]; out[1] = bv[1]; return; }
    out[0] = median3(a[0], bv[0], c[0]);
    out[1] = median3(a[1], bv[1], c[1]);
}

/* ---- tile decode: decode_intra.go, decode_inter.go ------------------------ */
typedef struct {
    rdec d;
    fbuf *rec, *ref, *golden;
    int32_t qdc, qac;
    int inter;
    int16_t *mv;
    int mvstride;
    orc_mbtrace *trace; /* this macroblock's trace record, or NULL */
    int blk;            /* block id being parsed, for the trace */
    int32_t levels[256], coeffs[256], recon[256], pred[256];
} tdec;

static void trace_levels(tdec *t, int n)
{
    if (!t->trace) return;
    for (int i = 0; i < n * n; i++)
        if (t->levels[i] && t->trace->ncoef < 384)
            t->trace->coef[t->trace->ncoef++] = (uint32_t)i | (uint32_t)t->blk << 8 | (uint32_t)t->levels[i] << 13;
}

static void dequant(const tdec *t, int n) /* quantizer.dequantize, quant.go:72-80 */
{
    int32_t *co = ((tdec *)t)->coeffs;
    for (int i = 0; i < n * n; i++) co[i] = t->levels[i] * (i == 0 ? t->qdc : t->qac);
}

/* reconBlock, decode_intra.go:170-187. */
static void recon_block(tdec *t, int plane_id, int tx, plane *p, int x0, int y0, int n, int mode,
                        int av_above, int av_left, int av_ar)
{
    if (mode < 0 || mode >= N_INTRA_MODES) { rfail(&t->d, "intra mode"); return; }
    memset(t->levels, 0, sizeof(int32_t) * n * n);
    decode_block(&t->d, plane_id, tx, t->levels, n);
    trace_levels(t, n);
    if (t->d.err) return;
    dequant(t, n);
    inverse_dct(t->coeffs, t->recon, n);
    predict_intra(p, x0, y0, n, mode, av_above, av_left, av_ar, t->pred);
    reconstruct(p, x0, y0, n, t->pred, t->recon);
}

/* decodeLumaMB, decode_intra.go:137-157. */
static void decode_luma_mb(tdec *t, int mbx, int mby, tbounds b)
{
    int tx = rdec_decode(&t->d, CTX_TXSIZE);
    if (tx < 0 || tx >= 3) { rfail(&t->d, "tx size"); return; }
    int n = 4 << tx;
    if (t->trace) t->trace->tx = tx;
    int tl = b.x0 * MBS, tt = b.y0 * MBS, tr = b.x1 * MBS, mbt = mby * MBS, mbr = (mbx + 1) * MBS;
    int bi = 0;
    for (int by = 0; by < MBS; by += n)
        for (int bx = 0; bx < MBS; bx += n) {
            int x0 = mbx * MBS + bx, y0 = mby * MBS + by;
            int mode = rdec_decode(&t->d, CTX_LUMAMODE);
            if (t->trace) t->trace->lmode[bi] = mode;
            t->blk = bi++;
            int ar = above_right_avail(x0, y0, n, tt, tr, mbt, mbr);
            recon_block(t, 0, tx, &t->rec->y, x0, y0, n, mode, y0 > tt, x0 > tl, ar);
        }
}

/* decodeChromaMB, decode_intra.go:159-167. */
static void decode_chroma_mb(tdec *t, int mbx, int mby, tbounds b)
{
    int tl = b.x0 * CMBS, tt = b.y0 * CMBS, x0 = mbx * CMBS, y0 = mby * CMBS;
    int mode = rdec_decode(&t->d, CTX_CHROMAMODE);
    if (t->trace) t->trace->cmode = mode;
    t->blk = 16;
    recon_block(t, 1, 0, &t->rec->cb, x0, y0, CMBS, mode, y0 > tt, x0 > tl, 0);
    t->blk = 17;
    recon_block(t, 1, 0, &t->rec->cr, x0, y0, CMBS, mode, y0 > tt, x0 > tl, 0);
}

/* reconResidual, decode_inter.go:140-152 (prediction already in t->pred). */
static void recon_residual(tdec *t, int plane_id, int tx, plane *p, int x0, int y0, int n)
{
    memset(t->levels, 0, sizeof(int32_t) * n * n);
    decode_block(&t->d, plane_id, tx, t->levels, n);
    trace_levels(t, n);
    if (t->d.err) return;
    dequant(t, n);
    inverse_dct(t->coeffs, t->recon, n);
    reconstruct(p, x0, y0, n, t->pred, t->recon);
}

/* reconInterChroma, decode_inter.go:123-136. */
static void recon_inter_chroma(tdec *t, int cx, int cy, int part, const int16_t (*mvs)[2], const plane *ref,
                               plane *rec, int coded)
{
    if (coded) {
        int32_t mc[64];
        for (int i = 0; i < part_count[part]; i++) {
            const int *rc = part_rect[part][i];
            int cpx = rc[0] / 2, cpy = rc[1] / 2;
            mc_chroma(ref, cx + cpx + (mvs[i][0] >> 3), cy + cpy + (mvs[i][1] >> 3), mvs[i][0] & 7,
                      mvs[i][1] & 7, rc[2] / 2, rc[3] / 2, mc, cpx, cpy, CMBS);
        }
        memcpy(t->pred, mc, sizeof mc);
        recon_residual(t, 1, 0, rec, cx, cy, CMBS);
        return;
    }
    for (int i = 0; i < part_count[part]; i++) {
        const int *rc = part_rect[part][i];
        int cpx = rc[0] / 2, cpy = rc[1] / 2;
        mc_chroma_byte(ref, cx + cpx + (mvs[i][0] >> 3), cy + cpy + (mvs[i][1] >> 3), mvs[i][0] & 7,
                       mvs[i][1] & 7, rc[2] / 2, rc[3] / 2, rec, cx + cpx, cy + cpy);
    }
}

/* reconInterMB, decode_inter.go:73-118. */
static void recon_inter_mb(tdec *t, const fbuf *ref, int mbx, int mby, int part, const int16_t (*mvs)[2])
{
    int cbp = rdec_decode(&t->d, CTX_CBP);
    if (t->trace) t->trace->cbp = cbp;
    if (cbp < 0 || cbp >= 8 || t->d.err) { rfail(&t->d, "cbp"); return; }
    if (cbp & CBP_LUMA) {
        int32_t mcy[256];
        for (int i = 0; i < part_count[part]; i++) {
            const int *rc = part_rect[part][i];
            mc_luma(&ref->y, mbx * MBS + rc[0] + (mvs[i][0] >> 2), mby * MBS + rc[1] + (mvs[i][1] >> 2),
                    mvs[i][0] & 3, mvs[i][1] & 3, rc[2], rc[3], mcy, rc[0], rc[1], MBS);
        }
        int tx = rdec_decode(&t->d, CTX_TXSIZE);
        if (tx < 0 || tx >= 3) { rfail(&t->d, "tx size"); return; }
        int n = 4 << tx;
        if (t->trace) t->trace->tx = tx;
        int bi = 0;
        for (int by = 0; by < MBS; by += n)
            for (int bx = 0; bx < MBS; bx += n) {
                t->blk = bi++;
                for (int r = 0; r < n; r++)
                    for (int c = 0; c < n; c++) t->pred[r * n + c] = mcy[(by + r) * MBS + bx + c];
                recon_residual(t, 0, tx, &t->rec->y, mbx * MBS + bx, mby * MBS + by, n);
                if (t->d.err) return;
            }
    } else {
        for (int i = 0; i < part_count[part]; i++) {
            const int *rc = part_rect[part][i];
            mc_luma_byte(&ref->y, mbx * MBS + rc[0] + (mvs[i][0] >> 2), mby * MBS + rc[1] + (mvs[i][1] >> 2),
                         mvs[i][0] & 3, mvs[i][1] & 3, rc[2], rc[3], &t->rec->y, mbx * MBS + rc[0],
                         mby * MBS + rc[1]);
        }
    }
    int cx = mbx * CMBS, cy = mby * CMBS;
    t->blk = 16;
    recon_inter_chroma(t, cx, cy, part, mvs, &ref->cb, &t->rec->cb, (cbp & CBP_CB) != 0);
    t->blk = 17;
    recon_inter_chroma(t, cx, cy, part, mvs, &ref->cr, &t->rec->cr, (cbp & CBP_CR) != 0);
}

/* decodeInterMB, decode_inter.go:8-62. */
static void decode_inter_mb(tdec *t, int mbx, int mby, tbounds b)
{
    int mode = rdec_decode(&t->d, CTX_MBMODE);
    int idx = mby * t->mvstride + mbx;
    int16_t pred[2];
    if (t->trace) t->trace->kind = mode;
    switch (mode) {
    case MB_SKIP: {
        predict_mv(t->mv, t->mvstride, mbx, mby, b, pred);
        if (t->trace) { t->trace->mv[0][0] = pred[0]; t->trace->mv[0][1] = pred[1]; }
        t->mv[2 * idx] = pred[0];
        t->mv[2 * idx + 1] = pred[1];
        /* reconMCMB, decode_inter.go:66-71 */
        const fbuf *r = t->ref;
        mc_luma_byte(&r->y, mbx * MBS + (pred[0] >> 2), mby * MBS + (pred[1] >> 2), pred[0] & 3, pred[1] & 3,
                     MBS, MBS, &t->rec->y, mbx * MBS, mby * MBS);
        int cx = mbx * CMBS, cy = mby * CMBS;
        mc_chroma_byte(&r->cb, cx + (pred[0] >> 3), cy + (pred[1] >> 3), pred[0] & 7, pred[1] & 7, CMBS, CMBS,
                       &t->rec->cb, cx, cy);
        mc_chroma_byte(&r->cr, cx + (pred[0] >> 3), cy + (pred[1] >> 3), pred[0] & 7, pred[1] & 7, CMBS, CMBS,
                       &t->rec->cr, cx, cy);
        break;
    }
    case MB_INTRA:
        decode_luma_mb(t, mbx, mby, b);
        decode_chroma_mb(t, mbx, mby, b);
        t->mv[2 * idx] = t->mv[2 * idx + 1] = 0;
        break;
    case MB_INTER: {
        int rs = rdec_decode(&t->d, CTX_REF);
        if (rs < 0 || rs >= 2) { rfail(&t->d, "ref select"); return; }
        const fbuf *ref = t->ref;
        predict_mv(t->mv, t->mvstride, mbx, mby, b, pred);
        if (rs == REF_GOLDEN) { ref = t->golden; pred[0] = pred[1] = 0; }
        int part = rdec_decode(&t->d, CTX_PART);
        if (part < 0 || part >= 4) { rfail(&t->d, "partition"); return; }
        int16_t mvs[4][2];
        for (int i = 0; i < part_count[part]; i++) {
            const int16_t *pp = i > 0 ? mvs[0] : pred;
            int px = pp[0], py = pp[1];
            int dx = zigzag_dec(class_value(&t->d, CTX_MVX));
            mvs[i][0] = clamp_mv(px + dx);
            int dy = zigzag_dec(class_value(&t->d, CTX_MVY));
            mvs[i][1] = clamp_mv(py + dy);
        }
        if (t->trace) {
            t->trace->refsel = rs;
            t->trace->part = part;
            for (int i = 0; i < part_count[part]; i++) { t->trace->mv[i][0] = mvs[i][0]; t->trace->mv[i][1] = mvs[i][1]; }
        }
        if (t->d.err) return;
        if (rs == REF_GOLDEN) { t->mv[2 * idx] = t->mv[2 * idx + 1] = 0; }
        else { t->mv[2 * idx] = mvs[0][0]; t->mv[2 * idx + 1] = mvs[0][1]; }
        recon_inter_mb(t, ref, mbx, mby, part, (const int16_t(*)[2])mvs);
        break;
    }
    default:
        rfail(&t->d, "macroblock mode");
    }
}

/* decodeTile, decode_intra.go:121-135. */
static void decode_tile(tdec *t, tbounds b, orc_mbtrace *trace, int mbc)
{
    for (int mby = b.y0; mby < b.y1; mby++)
        for (int mbx = b.x0; mbx < b.x1; mbx++) {
            t->trace = trace ? &trace[mby * mbc + mbx] : NULL;
            if (t->trace) { memset(t->trace, 0, sizeof *t->trace); t->trace->kind = 2; }
            if (t->inter) decode_inter_mb(t, mbx, mby, b);
            else { decode_luma_mb(t, mbx, mby, b); decode_chroma_mb(t, mbx, mby, b); }
            if (t->d.err) return;
        }
}

/* ---- codec: codec.go, frame.go --------------------------------------------- */
struct orc_codec {
    int ring_size, ready, next;
    int W, H, CW, CH, ystride, cstride;
    uint8_t **ring;
    fbuf bufs[3];
    fbuf *work, *ref, *golden;
    otable tables[N_CTX];
    oprev prev[N_CTX];
    int16_t *mv;
    int threads;
    uint64_t nsym;
    char msg[256];
    int trace_on;
    orc_mbtrace *trace;
    int ntrace;
    int in_tiles;  /* the last decode_coded_into failed in its tiles */
    int tile_fail; /* last decode: 0, 1 its own payload's tiles failed, 2 the override's */
};

static int fail_msg(orc_codec *c, int code, const char *m)
{
    snprintf(c->msg, sizeof c->msg, "%s", m);
    return code;
}

orc_codec *orc_codec_new(int ring)
{
    ensure_static();
    orc_codec *c = calloc(1, sizeof *c);
    c->ring_size = ring >= 2 ? ring : 3;
    c->threads = 1;
    return c;
}

void orc_codec_free(orc_codec *c)
{
    if (!c) return;
    if (c->ring) {
        for (int i = 0; i < c->ring_size; i++) free(c->ring[i]);
        free(c->ring);
    }
    for (int i = 0; i < 3; i++) { free(c->bufs[i].y.data); free(c->bufs[i].cb.data); free(c->bufs[i].cr.data); }
    free(c->mv);
    free(c->trace);
    free(c);
}

void orc_codec_enable_trace(orc_codec *c, int on) { c->trace_on = on; }
const orc_mbtrace *orc_codec_trace(const orc_codec *c, int *count)
{
    *count = c->ntrace;
    return c->trace;
}

void orc_codec_flush(orc_codec *c) { c->next = 0; }
const char *orc_codec_last_error(const orc_codec *c) { return c->msg; }
void orc_codec_set_threads(orc_codec *c, int n) { c->threads = n < 1 ? 1 : n; }
uint64_t orc_codec_symbols(const orc_codec *c) { return c->nsym; }

int orc_codec_golden(const orc_codec *c, const uint8_t **y, const uint8_t **cb, const uint8_t **cr, int *ypw,
                     int *yph)
{
    if (!c->ready) return -1;
    *y = c->golden->y.data; *cb = c->golden->cb.data; *cr = c->golden->cr.data;
    *ypw = c->golden->y.w; *yph = c->golden->y.h;
    return 0;
}

static void new_fbuf(fbuf *f, int w, int h) /* newFrameBuf, frame.go:25-37 */
{
    f->mbc = (w + MBS - 1) / MBS;
    f->mbr = (h + MBS - 1) / MBS;
    int pw = f->mbc * MBS, ph = f->mbr * MBS, cw = f->mbc * CMBS, ch = f->mbr * CMBS;
    f->y = (plane){calloc((size_t)pw * ph, 1), pw, pw, ph};
    f->cb = (plane){calloc((size_t)cw * ch, 1), cw, cw, ch};
    f->cr = (plane){calloc((size_t)cw * ch, 1), cw, cw, ch};
}

/* configure, codec.go:129-153. */
static int configure(orc_codec *c, int w, int h)
{
    if ((long)w * h > MAX_DECODE_PIXELS) return fail_msg(c, ORC_ERR_UNSUPPORTED, "frame exceeds the decoder's pixel bound");
    if (c->ready) {
        if (w != c->W || h != c->H) return fail_msg(c, ORC_ERR_DIMENSIONS, "key frame dimensions differ from the stream");
        return 0;
    }
    c->W = w; c->H = h; c->CW = (w + 1) / 2; c->CH = (h + 1) / 2;
    c->ystride = (w + 63) & ~63;
    c->cstride = (c->CW + 63) & ~63;
    c->ring = calloc((size_t)c->ring_size, sizeof *c->ring);
    size_t ys = (size_t)c->ystride * h, cs = (size_t)c->cstride * c->CH;
    for (int i = 0; i < c->ring_size; i++) c->ring[i] = calloc(ys + 2 * cs, 1);
    for (int i = 0; i < 3; i++) new_fbuf(&c->bufs[i], w, h);
    c->work = &c->bufs[0]; c->ref = &c->bufs[1]; c->golden = &c->bufs[2];
    c->mv = calloc((size_t)c->work->mbc * c->work->mbr * 2, sizeof(int16_t));
    c->next = 0;
    c->ready = 1;
    return 0;
}

typedef struct {
    orc_codec *c;
    fbuf *rec, *ref, *golden;
    int inter, ntiles;
    int32_t qdc, qac;
    const uint8_t **tile_ptr;
    const size_t *tile_len;
    const int *tb; /* 4 ints per tile */
    int *errs;
    const char **msgs;
    uint64_t *nsym;
    orc_mbtrace *trace;
    atomic_int next;
} frame_job;

static void run_tile(frame_job *j, int i)
{
    tdec *t = malloc(sizeof *t);
    memset(t, 0, offsetof(tdec, levels));
    t->rec = j->rec; t->ref = j->ref; t->golden = j->golden;
    t->qdc = j->qdc; t->qac = j->qac; t->inter = j->inter;
    t->mv = j->c->mv; t->mvstride = j->rec->mbc;
    if (rdec_reset(&t->d, j->tile_ptr[i], j->tile_len[i], j->c->tables) < 0) {
        j->errs[i] = ORC_ERR_CORRUPT;
        j->msgs[i] = "tile stream too short to initialize rANS";
        free(t);
        return;
    }
    tbounds b = {j->tb[4 * i], j->tb[4 * i + 1], j->tb[4 * i + 2], j->tb[4 * i + 3]};
    decode_tile(t, b, j->trace, j->rec->mbc);
    j->errs[i] = t->d.err;
    j->msgs[i] = t->d.msg;
    j->nsym[i] = t->d.nsym;
    free(t);
}

static void *tile_worker(void *arg)
{
    frame_job *j = arg;
    for (;;) {
        int i = atomic_fetch_add(&j->next, 1);
        if (i >= j->ntiles) break;
        run_tile(j, i);
    }
    return NULL;
}

/* decodeState.decodeCoded (decode_intra.go:55-94) with parsePayload (frame.go:242-274). */
static int decode_coded_into(orc_codec *c, fbuf *rec, fbuf *ref, fbuf *golden, const uint8_t *body, size_t len,
                             int inter, int key)
{
    /* parseCodedHeader, frame.go:197-220 */
    if (len < 1) return fail_msg(c, ORC_ERR_CORRUPT, "coded header truncated");
    int qp = body[0];
    size_t pos = 1;
    uint64_t cols, rows, tlen;
    int n = uvarint(body + pos, len - pos, &cols);
    if (n <= 0) return fail_msg(c, ORC_ERR_CORRUPT, "coded header tile columns");
    pos += (size_t)n;
    n = uvarint(body + pos, len - pos, &rows);
    if (n <= 0) return fail_msg(c, ORC_ERR_CORRUPT, "coded header tile rows");
    pos += (size_t)n;
    if (qp > MAX_QP || cols == 0 || rows == 0 || cols > 65535 || rows > 65535)
        return fail_msg(c, ORC_ERR_CORRUPT, "coded header values out of range");
    n = uvarint(body + pos, len - pos, &tlen);
    /* Go compares pos+n+int(tablesLen) > len; lengths >= 2^63 wrap there and
     * then panic on the slice, so any length past the body is corrupt here. */
    if (n <= 0 || tlen > (uint64_t)(len - pos - (size_t)n)) return fail_msg(c, ORC_ERR_CORRUPT, "bad table length");
    pos += (size_t)n;
    const char *m = NULL;
    long consumed = parse_tables(c->tables, c->prev, body + pos, (size_t)tlen, key, &m);
    if (consumed < 0) return fail_msg(c, ORC_ERR_CORRUPT, m);
    if ((uint64_t)consumed != tlen) return fail_msg(c, ORC_ERR_CORRUPT, "trailing table bytes");
    pos += (size_t)tlen;

    uint64_t want = cols * rows;
    size_t cap = 16, count = 0;
    const uint8_t **tp = malloc(cap * sizeof *tp);
    size_t *tl = malloc(cap * sizeof *tl);
    while (count < want) {
        uint64_t l;
        n = uvarint(body + pos, len - pos, &l);
        if (n <= 0 || l > (uint64_t)(len - pos - (size_t)n)) {
            free(tp); free(tl);
            return fail_msg(c, ORC_ERR_CORRUPT, "bad tile length");
        }
        pos += (size_t)n;
        if (count == cap) { cap *= 2; tp = realloc(tp, cap * sizeof *tp); tl = realloc(tl, cap * sizeof *tl); }
        tp[count] = body + pos;
        tl[count] = (size_t)l;
        count++;
        pos += (size_t)l;
    }
    /* newTileGrid, frame.go:87-111 */
    int mbc = rec->mbc, mbr = rec->mbr;
    int gc = (int)cols < mbc ? (int)cols : mbc, gr = (int)rows < mbr ? (int)rows : mbr;
    if ((size_t)gc * gr != count) { free(tp); free(tl); return fail_msg(c, ORC_ERR_CORRUPT, "ebiv: corrupt stream"); }
    int nt = (int)count;
    int *tb = malloc(sizeof(int) * 4 * nt);
    for (int i = 0; i < nt; i++) { /* tileGrid.bounds, frame.go:170-177 */
        int tc = i % gc, tr = i / gc;
        tb[4 * i] = mbc * tc / gc; tb[4 * i + 1] = mbc * (tc + 1) / gc;
        tb[4 * i + 2] = mbr * tr / gr; tb[4 * i + 3] = mbr * (tr + 1) / gr;
    }
    int q = qp; /* newQuantizer, quant.go:38-51 */
    frame_job j = {0};
    j.c = c; j.rec = rec; j.ref = ref; j.golden = golden; j.inter = inter; j.ntiles = nt;
    j.qac = qstep[q]; j.qdc = q >= 6 ? qstep[q - 6] : qstep[q];
    j.tile_ptr = tp; j.tile_len = tl; j.tb = tb;
    j.errs = calloc((size_t)nt, sizeof(int));
    j.msgs = calloc((size_t)nt, sizeof(char *));
    j.nsym = calloc((size_t)nt, sizeof(uint64_t));
    atomic_init(&j.next, 0);
    memset(c->mv, 0, sizeof(int16_t) * 2 * (size_t)mbc * mbr); /* resetMV */
    j.trace = NULL;
    if (c->trace_on && rec == c->work) {
        if (!c->trace) { c->trace = calloc((size_t)mbc * mbr, sizeof(orc_mbtrace)); c->ntrace = mbc * mbr; }
        for (int i = 0; i < mbc * mbr; i++) c->trace[i].kind = -1;
        j.trace = c->trace;
    }
    int workers = c->threads < nt ? c->threads : nt;
    if (workers <= 1) {
        for (int i = 0; i < nt; i++) run_tile(&j, i);
    } else {
        pthread_t th[256];
        if (workers > 256) workers = 256;
        for (int w = 0; w < workers; w++) pthread_create(&th[w], NULL, tile_worker, &j);
        for (int w = 0; w < workers; w++) pthread_join(th[w], NULL);
    }
    int rc = 0;
    for (int i = 0; i < nt; i++) {
        c->nsym += j.nsym[i];
        if (!rc && j.errs[i]) rc = fail_msg(c, j.errs[i], j.msgs[i] ? j.msgs[i] : "corrupt tile");
    }
    c->in_tiles = rc != 0;
    free(j.errs); free(j.msgs); free(j.nsym); free(tb); free(tp); free(tl);
    return rc;
}

static void store_plane(uint8_t *dst, int ds, const plane *src, int w, int h) /* storePlane, frame.go:73-77 */
{
    for (int y = 0; y < h; y++) memcpy(dst + (size_t)y * ds, src->data + (size_t)y * src->stride, (size_t)w);
}

static void fill_frame(orc_codec *c, uint8_t *img, orc_frame *out)
{
    size_t ys = (size_t)c->ystride * c->H, cs = (size_t)c->cstride * c->CH;
    out->y = img; out->cb = img + ys; out->cr = img + ys + cs;
    out->y_stride = c->ystride; out->c_stride = c->cstride;
    out->width = c->W; out->height = c->H; out->cwidth = c->CW; out->cheight = c->CH;
}

int orc_codec_last_tile_failure(const orc_codec *c) { return c->tile_fail; }

int orc_codec_decode(orc_codec *c, const uint8_t *data, size_t len, orc_frame *out)
{
    c->tile_fail = 0;
    c->in_tiles = 0;
    /* parseFrameHeader, format.go:376-409 */
    if (len < 1) return fail_msg(c, ORC_ERR_CORRUPT, "empty frame payload");
    int type = data[0] & 3, coding = (data[0] >> 2) & 7, gov = (data[0] >> 5) & 1;
    if (type != 0 && type != 1) return fail_msg(c, ORC_ERR_CORRUPT, "frame type");
    if (gov && type != 0) return fail_msg(c, ORC_ERR_CORRUPT, "golden override on a non-key frame");
    const uint8_t *body = data + 1;
    size_t blen = len - 1;
    if (type == 0) {
        if (len < 6) return fail_msg(c, ORC_ERR_CORRUPT, "key frame header truncated");
        int w = data[1] | data[2] << 8, h = data[3] | data[4] << 8;
        if (w == 0 || h == 0) return fail_msg(c, ORC_ERR_CORRUPT, "key frame dimensions");
        if ((data[5] >> 4) != 0) return fail_msg(c, ORC_ERR_UNSUPPORTED, "chroma format");
        if ((data[5] & 15) != 0) return fail_msg(c, ORC_ERR_UNSUPPORTED, "bit depth code");
        body = data + 6;
        blen = len - 6;
        int rc = configure(c, w, h);
        if (rc) return rc;
    } else if (!c->ready) {
        return fail_msg(c, ORC_ERR_NOKEYFRAME, "inter frame before any key frame");
    }
    uint8_t *img = c->ring[c->next];
    if (++c->next == c->ring_size) c->next = 0;
    size_t ys = (size_t)c->ystride * c->H, cs = (size_t)c->cstride * c->CH;

    if (coding == 0) { /* decodeRaw, codec.go:195-207 */
        size_t want = (size_t)c->W * c->H + 2 * (size_t)c->CW * c->CH;
        if (blen != want) return fail_msg(c, ORC_ERR_CORRUPT, "raw frame size");
        for (int y = 0; y < c->H; y++) memcpy(img + (size_t)y * c->ystride, body + (size_t)y * c->W, (size_t)c->W);
        const uint8_t *cb = body + (size_t)c->W * c->H, *cr = cb + (size_t)c->CW * c->CH;
        for (int y = 0; y < c->CH; y++) {
            memcpy(img + ys + (size_t)y * c->cstride, cb + (size_t)y * c->CW, (size_t)c->CW);
            memcpy(img + ys + cs + (size_t)y * c->cstride, cr + (size_t)y * c->CW, (size_t)c->CW);
        }
        fill_frame(c, img, out);
        return 0;
    }
    if (coding != 1 && coding != 2) return fail_msg(c, ORC_ERR_UNSUPPORTED, "coding mode");

    /* Codec.decodeCoded, codec.go:162-192 */
    int inter = coding == 2, key = type == 0;
    const uint8_t *ov = NULL;
    size_t ovlen = 0;
    if (gov) {
        uint64_t kl;
        int n = uvarint(body, blen, &kl);
        if (n <= 0 || kl > (uint64_t)(blen - (size_t)n)) return fail_msg(c, ORC_ERR_CORRUPT, "golden override split");
        ov = body + n + kl;
        ovlen = blen - (size_t)n - (size_t)kl;
        body = body + n;
        blen = (size_t)kl;
    }
    int rc = decode_coded_into(c, c->work, c->ref, c->golden, body, blen, inter, key);
    if (rc) {
        c->tile_fail = c->in_tiles ? 1 : 0;
        return rc;
    }
    if (key) {
        if (ov) {
            c->in_tiles = 0;
            rc = decode_coded_into(c, c->golden, c->work, c->work, ov, ovlen, 1, 0);
            if (rc) {
                c->tile_fail = c->in_tiles ? 2 : 0;
                return rc;
            }
        } else {
            fbuf *g = c->golden, *w = c->work; /* copyFrom, frame.go:42-46 */
            memcpy(g->y.data, w->y.data, (size_t)w->y.stride * w->y.h);
            memcpy(g->cb.data, w->cb.data, (size_t)w->cb.stride * w->cb.h);
            memcpy(g->cr.data, w->cr.data, (size_t)w->cr.stride * w->cr.h);
        }
    }
    store_plane(img, c->ystride, &c->work->y, c->W, c->H);
    store_plane(img + ys, c->cstride, &c->work->cb, c->CW, c->CH);
    store_plane(img + ys + cs, c->cstride, &c->work->cr, c->CW, c->CH);
    fbuf *t = c->work; c->work = c->ref; c->ref = t;
    fill_frame(c, img, out);
    return 0;
}

/* convertYCbCr420ToRGBA, frame.go:8-67. */
void orc_ycbcr_to_rgba(const uint8_t *y, int ys, const uint8_t *cb, const uint8_t *cr, int cs, int w, int h,
                       uint8_t *dst)
{
    size_t di = 0;
    for (int row = 0; row < h; row++) {
        const uint8_t *yr = y + (size_t)row * ys;
        const uint8_t *cbr = cb + (size_t)(row / 2) * cs, *crr = cr + (size_t)(row / 2) * cs;
        for (int col = 0; col < w; col++) {
            int32_t yy = (int32_t)yr[col] * 0x10101;
            int32_t b1 = (int32_t)cbr[col / 2] - 128, r1 = (int32_t)crr[col / 2] - 128;
            int32_t r = yy + 91881 * r1, g = yy - 22554 * b1 - 46802 * r1, b = yy + 116130 * b1;
            r = r < 0 ? 0 : r > 0xff0000 ? 0xff0000 : r;
            g = g < 0 ? 0 : g > 0xff0000 ? 0xff0000 : g;
            b = b < 0 ? 0 : b > 0xff0000 ? 0xff0000 : b;
            dst[di] = (uint8_t)(r >> 16); dst[di + 1] = (uint8_t)(g >> 16);
            dst[di + 2] = (uint8_t)(b >> 16); dst[di + 3] = 0xff;
            di += 4;
        }
    }
}

/* ---- unit-level entry points ---------------------------------------------- */
int orc_inverse_dct(const int32_t *src, int32_t *dst, int n)
{
    ensure_static();
    if (n != 4 && n != 8 && n != 16) return -1;
    inverse_dct(src, dst, n);
    return 0;
}
int32_t orc_dct_coef(int n, int k, int i) { ensure_static(); return dct[size_idx(n)][k * n + i]; }
int32_t orc_quant_step(int qp) { ensure_static(); return qstep[qp < 0 ? 0 : qp > MAX_QP ? MAX_QP : qp]; }
int orc_zigzag(int n, int i) { ensure_static(); return zz[size_idx(n)][i]; }

int orc_predict_intra(const uint8_t *p, int stride, int w, int h, int x0, int y0, int n, int mode, int av_above,
                      int av_left, int av_ar, int32_t *pred)
{
    ensure_static();
    if (mode < 0 || mode >= N_INTRA_MODES) return -1;
    plane pl = {(uint8_t *)p, stride, w, h};
    predict_intra(&pl, x0, y0, n, mode, av_above, av_left, av_ar, pred);
    return 0;
}
int orc_above_right_available(int x0, int y0, int n, int tt, int tr, int mbt, int mbr)
{
    return above_right_avail(x0, y0, n, tt, tr, mbt, mbr);
}
void orc_mc_luma(const uint8_t *p, int stride, int w, int h, int sx, int sy, int fx, int fy, int pw, int ph,
                 int32_t *out, int os)
{
    plane pl = {(uint8_t *)p, stride, w, h};
    mc_luma(&pl, sx, sy, fx, fy, pw, ph, out, 0, 0, os);
}
void orc_mc_chroma(const uint8_t *p, int stride, int w, int h, int sx, int sy, int fx, int fy, int pw, int ph,
                   int32_t *out, int os)
{
    plane pl = {(uint8_t *)p, stride, w, h};
    mc_chroma(&pl, sx, sy, fx, fy, pw, ph, out, 0, 0, os);
}
void orc_predict_mv(const int16_t *mv, int stride, int mbx, int mby, int x0, int x1, int y0, int y1, int16_t *out)
{
    tbounds b = {x0, x1, y0, y1};
    predict_mv(mv, stride, mbx, mby, b, out);
}

static int setup_tables(otable *tables, const uint8_t *blob, size_t len)
{
    oprev prev[N_CTX];
    memset(prev, 0, sizeof prev);
    const char *m;
    long k = parse_tables(tables, prev, blob, len, 1, &m);
    return (k < 0 || (size_t)k != len) ? -1 : 0;
}

int orc_rans_decode_symbols(const uint8_t *blob, size_t blen, const uint8_t *stream, size_t slen, const int *ctxs,
                            int count, int *out)
{
    ensure_static();
    otable *tables = malloc(sizeof(otable) * N_CTX);
    if (setup_tables(tables, blob, blen) < 0) { free(tables); return ORC_ERR_CORRUPT; }
    rdec d;
    if (rdec_reset(&d, stream, slen, tables) < 0) { free(tables); return ORC_ERR_CORRUPT; }
    for (int i = 0; i < count; i++) out[i] = rdec_decode(&d, ctxs[i]);
    free(tables);
    return d.err;
}

int orc_decode_blocks(const uint8_t *blob, size_t blen, const uint8_t *stream, size_t slen, int plane_id, int tx,
                      int n, int count, int32_t *levels)
{
    ensure_static();
    otable *tables = malloc(sizeof(otable) * N_CTX);
    if (setup_tables(tables, blob, blen) < 0) { free(tables); return ORC_ERR_CORRUPT; }
    rdec d;
    if (rdec_reset(&d, stream, slen, tables) < 0) { free(tables); return ORC_ERR_CORRUPT; }
    memset(levels, 0, sizeof(int32_t) * (size_t)count * n * n);
    for (int b = 0; b < count; b++) decode_block(&d, plane_id, tx, levels + (size_t)b * n * n, n);
    free(tables);
    return d.err;
}
