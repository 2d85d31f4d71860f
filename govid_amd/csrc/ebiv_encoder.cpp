// ebiv_encoder.cpp — EBIV stream encoder (host, C++17).
//
// Restates govid's offline encoder (ebiv/encoder.go, ebiv/encode_intra.go,
// ebiv/encode_inter.go and the encode halves of rans.go, residual.go, quant.go,
// transform.go, mv.go, writer.go) so that streams exercising every decode
// syntax element can be produced where Go is absent.  Pass 1 (dead-zone
// quantisation priced with estimateBits) is what WithFastEncode runs; the
// second, real-cost RDOQ pass (quant.go:102-199) is run when fast == 0.
// Its reconstruction is computed by its own integer paths, independent of
// the decoder, so "encoder reconstruction == decoded output" is a real check.
//
// C ABI at the bottom (ebivenc_*) for the Python harness.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "ebiv_format.h"

namespace ebiv {
namespace {

// ---- static tables ------------------------------------------------------------
struct Tables {
  int zz[3][256];
  int32_t dct[3][256];
  int32_t qstep[kMaxQP + 1];
  Tables() {
    for (int s = 0; s < 3; s++) {  // buildZigzag, scan.go:17-35
      int n = 4 << s, k = 0;
      for (int d = 0; d < 2 * n - 1; d++) {
        if (d % 2 == 0) {
          for (int y = std::min(d, n - 1); y >= 0 && d - y < n; y--) zz[s][k++] = y * n + d - y;
        } else {
          for (int x = std::min(d, n - 1); x >= 0 && d - x < n; x--) zz[s][k++] = (d - x) * n + x;
        }
      }
      for (int kk = 0; kk < n; kk++) {  // buildDCTMatrix, transform.go:27-40
        double alpha = kk == 0 ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
        for (int i = 0; i < n; i++)
          dct[s][kk * n + i] = static_cast<int32_t>(std::round(alpha * std::cos((2.0 * i + 1.0) * kk * M_PI / (2.0 * n)) * 8192.0));
      }
    }
    for (int qp = 0; qp <= kMaxQP; qp++) {  // quant.go:19-29
      double st = std::round(std::pow(2.0, qp / 6.0));
      qstep[qp] = static_cast<int32_t>(st < 1 ? 1 : st);
    }
  }
};
const Tables& T() {
  static Tables t;
  return t;
}
inline int sidx(int n) { return n == 4 ? 0 : n == 8 ? 1 : 2; }
inline uint8_t clamp_byte(int32_t v) { return v < 0 ? 0 : v > 255 ? 255 : static_cast<uint8_t>(v); }
inline int32_t abs32(int32_t v) { return v < 0 ? -v : v; }
inline int bits_len(unsigned v) { return v == 0 ? 0 : 32 - __builtin_clz(v); }

// ---- planes ----------------------------------------------------------------------
struct Plane {
  std::vector<uint8_t> d;
  int stride = 0, w = 0, h = 0;
  uint8_t at(int x, int y) const { return d[static_cast<size_t>(y) * stride + x]; }
  void set(int x, int y, uint8_t v) { d[static_cast<size_t>(y) * stride + x] = v; }
};
struct FrameBuf {
  Plane y, cb, cr;
  int mbc = 0, mbr = 0;
  FrameBuf(int W, int H) {  // newFrameBuf, frame.go:25-37
    mbc = (W + kMB - 1) / kMB;
    mbr = (H + kMB - 1) / kMB;
    auto mk = [](Plane& p, int w, int h) { p.d.assign(static_cast<size_t>(w) * h, 0); p.stride = p.w = w; p.h = h; };
    mk(y, mbc * kMB, mbr * kMB);
    mk(cb, mbc * kCMB, mbr * kCMB);
    mk(cr, mbc * kCMB, mbr * kCMB);
  }
};
using FrameBufP = std::shared_ptr<FrameBuf>;

struct SrcImage {  // packed visible 4:2:0 planes
  int W, H, CW, CH;
  std::vector<uint8_t> y, cb, cr;
};

// loadPlane, frame.go:56-64: replicate the visible edge into the padding.
void load_plane(Plane& dst, const std::vector<uint8_t>& src, int vw, int vh) {
  for (int y = 0; y < dst.h; y++) {
    int sy = std::min(y, vh - 1);
    for (int x = 0; x < dst.w; x++) dst.d[static_cast<size_t>(y) * dst.stride + x] = src[static_cast<size_t>(sy) * vw + std::min(x, vw - 1)];
  }
}

struct MV { int16_t x = 0, y = 0; };
inline bool operator==(MV a, MV b) { return a.x == b.x && a.y == b.y; }
inline int16_t clamp_mv(int v) { return static_cast<int16_t>(std::clamp(v, -kMVClamp, kMVClamp)); }

struct Tok { uint16_t ctx, sym; };
static_assert(sizeof(Tok) == 4, "token log exports (ctx, sym) as u16 pairs");
struct TileStream {
  std::vector<Tok> toks;
  void put(int ctx, int sym) { toks.push_back({static_cast<uint16_t>(ctx), static_cast<uint16_t>(sym)}); }
};

// ---- transforms (transform.go) --------------------------------------------------
void forward_dct(const int32_t* src, int32_t* dst, int n) {
  const int32_t* c = T().dct[sidx(n)];
  int64_t acc[256] = {0};
  int32_t tmp[256];
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++) {
      int64_t cki = c[k * n + i];
      for (int j = 0; j < n; j++) acc[k * n + j] += cki * src[i * n + j];
    }
  for (int k = 0; k < n * n; k++) tmp[k] = static_cast<int32_t>((acc[k] + 4096) >> 13);
  for (int k = 0; k < n; k++)
    for (int l = 0; l < n; l++) {
      int64_t a = 0;
      for (int j = 0; j < n; j++) a += static_cast<int64_t>(c[l * n + j]) * tmp[k * n + j];
      dst[k * n + l] = static_cast<int32_t>((a + 4096) >> 13);
    }
}
// Full separable inverse: zero terms add nothing, so this equals the sparse
// form of transform.go:129-194 exactly.
void inverse_dct(const int32_t* src, int32_t* dst, int n) {
  const int32_t* c = T().dct[sidx(n)];
  int32_t tmp[256];
  for (int i = 0; i < n; i++)
    for (int l = 0; l < n; l++) {
      int64_t a = 0;
      for (int k = 0; k < n; k++) a += static_cast<int64_t>(c[k * n + i]) * src[k * n + l];
      tmp[i * n + l] = static_cast<int32_t>((a + 4096) >> 13);
    }
  for (int i = 0; i < n; i++)
    for (int m = 0; m < n; m++) {
      int64_t a = 0;
      for (int l = 0; l < n; l++) a += static_cast<int64_t>(c[l * n + m]) * tmp[i * n + l];
      dst[i * n + m] = static_cast<int32_t>((a + 4096) >> 13);
    }
}
void fwht(int32_t* a, int n, int step) {
  for (int span = 1; span < n; span <<= 1)
    for (int i = 0; i < n; i += span << 1)
      for (int j = i; j < i + span; j++) {
        int32_t x = a[j * step], y = a[(j + span) * step];
        a[j * step] = x + y;
        a[(j + span) * step] = x - y;
      }
}
int64_t satd(const int32_t* res, int n) {  // transform.go:86-107
  int32_t t[256];
  std::memcpy(t, res, sizeof(int32_t) * n * n);
  for (int r = 0; r < n; r++) fwht(t + r * n, n, 1);
  for (int c = 0; c < n; c++) fwht(t + c, n, n);
  int64_t s = 0;
  for (int i = 0; i < n * n; i++) s += abs32(t[i]);
  return s;
}

// ---- quantiser (quant.go) ---------------------------------------------------------
struct Quant {
  int32_t dc, ac;
  explicit Quant(int qp) {
    qp = std::clamp(qp, 0, kMaxQP);
    ac = T().qstep[qp];
    dc = qp >= 6 ? T().qstep[qp - 6] : ac;
  }
  void quantize(const int32_t* co, int32_t* lv, int n) const {  // quant.go:55-70
    for (int i = 0; i < n * n; i++) {
      int32_t step = i == 0 ? dc : ac, c = co[i];
      lv[i] = c >= 0 ? (c * 8 + 3 * step) / (step * 8) : -((-c * 8 + 3 * step) / (step * 8));
    }
  }
  void dequantize(const int32_t* lv, int32_t* co, int n) const {
    for (int i = 0; i < n * n; i++) co[i] = lv[i] * (i == 0 ? dc : ac);
  }
};

using CostModel = std::vector<std::array<double, kMaxAlphabet>>;

// Go's math.Log2: Frexp, then Log(frac)*(1/Ln2) + exp (exact for powers of 2).
double go_log2(double x) {
  int e;
  double frac = std::frexp(x, &e);
  if (frac == 0.5) return static_cast<double>(e - 1);
  return std::log(frac) * (1.0 / M_LN2) + static_cast<double>(e);
}

double coeff_rate(int m, int ctx, const CostModel& cost) {  // quant.go:202-208
  if (m <= 4) return cost[ctx][m];
  unsigned v = static_cast<unsigned>(m - 5);
  int bl = bits_len(v);
  return cost[ctx][kTEscape] + cost[kCtxEscClass][bl] + static_cast<double>(std::max(0, bl - 1));
}

// quantizeRDOQ, quant.go:102-199.
void quantize_rdoq(const Quant& q, const int32_t* coeffs, int32_t* levels, int n, int plane, int tx, double lambda,
                   const CostModel& cost) {
  const int* scan = T().zz[sidx(n)];
  int total = n * n;
  double sign_bits = (cost[kCtxSign][0] + cost[kCtxSign][1]) * 0.5;
  int ctx_scan[256];
  int32_t lvl_scan[256];
  double d_scan[256], rnz_scan[256];
  int prev = 0;
  for (int si = 0; si < total; si++) {
    int nat = scan[si];
    double step = nat == 0 ? static_cast<double>(q.dc) : static_cast<double>(q.ac);
    int32_t c = coeffs[nat];
    double ac = static_cast<double>(abs32(c));
    int ctx = token_ctx(plane, tx, band_of(si), prev);
    ctx_scan[si] = ctx;
    int l0 = static_cast<int>((ac + step / 2) / step);
    int best_mag = 0;
    double best_d = ac * ac, best_cost = best_d, best_r = 0.0;
    for (int cand : {l0 - 1, l0}) {
      if (cand < 1) continue;
      double d = ac - static_cast<double>(cand) * step;
      d *= d;
      double r = coeff_rate(cand, ctx, cost) + sign_bits;
      double cst = d + lambda * r;
      if (cst < best_cost) { best_cost = cst; best_mag = cand; best_d = d; best_r = r; }
    }
    d_scan[si] = best_d;
    rnz_scan[si] = best_r;
    if (best_mag > 0) {
      lvl_scan[si] = c < 0 ? -best_mag : best_mag;
      prev = class_for_mag(best_mag);
    } else {
      lvl_scan[si] = 0;
      prev = 0;
    }
  }
  double suffix[257];
  suffix[total] = 0;
  for (int i = total - 1; i >= 0; i--) {
    double c = static_cast<double>(coeffs[scan[i]]);
    suffix[i] = suffix[i + 1] + c * c;
  }
  int last_nz = -1;
  for (int si = 0; si < total; si++) if (lvl_scan[si] != 0) last_nz = si;
  int best_t = -1;
  double best_cost = suffix[0] + lambda * cost[ctx_scan[0]][kTEOB];
  double pref_d = 0.0, pref_r = 0.0;
  for (int t = 0; t <= last_nz; t++) {
    pref_d += d_scan[t];
    pref_r += lvl_scan[t] != 0 ? rnz_scan[t] : cost[ctx_scan[t]][kTZero];
    double eob = t < total - 1 ? cost[ctx_scan[t + 1]][kTEOB] : 0.0;
    double cst = pref_d + suffix[t + 1] + lambda * (pref_r + eob);
    if (cst < best_cost) { best_cost = cst; best_t = t; }
  }
  for (int si = 0; si < total; si++) levels[scan[si]] = si <= best_t ? lvl_scan[si] : 0;
}

// applySignDataHiding, residual.go:134-180.
void apply_sdh(int32_t* lv, int n) {
  const int* scan = T().zz[sidx(n)];
  int total = n * n, first = -1, last = -1;
  int32_t sum = 0;
  for (int i = 0; i < total; i++) {
    int32_t v = abs32(lv[scan[i]]);
    if (v) { if (first < 0) first = i; last = i; sum += v; }
  }
  if (first < 0 || last - first < kSDHMinSpan) return;
  int sign = lv[scan[first]] < 0 ? 1 : 0;
  if ((sum & 1) == sign) return;
  for (int i = last; i >= first; i--) {
    int32_t v = lv[scan[i]];
    if (abs32(v) > 1) { lv[scan[i]] = v > 0 ? v - 1 : v + 1; return; }
  }
  for (int i = first + 1; i < last; i++)
    if (lv[scan[i]] == 0) { lv[scan[i]] = 1; return; }
  if (lv[scan[first + 1]] > 0) lv[scan[first + 1]]++; else lv[scan[first + 1]]--;
}

// putRawBits / putClassValue / putMVComponent / putEscapeValue, mv.go:67-123.
void put_class_value(TileStream& s, int ctx, unsigned v) {
  int cls = bits_len(v);
  s.put(ctx, cls);
  if (cls > 1) {
    unsigned r = v - (1u << (cls - 1));
    for (int i = cls - 2; i >= 0; i--) s.put(kCtxBypass, (r >> i) & 1);
  }
}
inline unsigned zigzag_enc(int v) { return static_cast<unsigned>((v << 1) ^ (v >> 31)); }

// encodeBlock, residual.go:30-75.  Returns true for an all-zero block.
bool encode_block(TileStream& s, int plane, int tx, const int32_t* lv, int n) {
  const int* scan = T().zz[sidx(n)];
  int total = n * n, first = -1, last = -1;
  for (int i = 0; i < total; i++)
    if (lv[scan[i]] != 0) { if (first < 0) first = i; last = i; }
  int prev = 0;
  for (int i = 0; i <= last; i++) {
    int32_t c = lv[scan[i]];
    int ctx = token_ctx(plane, tx, band_of(i), prev);
    int mag = abs32(c);
    if (mag == 0) { s.put(ctx, kTZero); prev = 0; }
    else if (mag <= 4) { s.put(ctx, mag); prev = class_for_mag(mag); }
    else { s.put(ctx, kTEscape); put_class_value(s, kCtxEscClass, static_cast<unsigned>(mag - 5)); prev = 2; }
    if (mag != 0 && i != first) s.put(kCtxSign, c < 0 ? 1 : 0);
  }
  if (last < total - 1) s.put(token_ctx(plane, tx, band_of(last + 1), prev), kTEOB);
  if (first >= 0 && last - first < kSDHMinSpan) s.put(kCtxSign, lv[scan[first]] < 0 ? 1 : 0);
  return last < 0;
}

// ---- intra prediction (predict.go) ---------------------------------------------
bool above_right_avail(int x0, int y0, int n, int tt, int tr, int mbt, int mbr) {
  if (y0 <= tt || x0 + 2 * n > tr) return false;
  if (y0 == mbt) return true;
  return x0 + 2 * n <= mbr;
}
void predict_intra(const Plane& p, int x0, int y0, int n, int mode, bool av_above, bool av_left, bool av_ar,
                   int32_t* pred) {
  int32_t ref[4 * kMB + 1] = {0};
  bool avail[4 * kMB + 1] = {false};
  int corner = 2 * n;
  for (int i = 0; i < 2 * n; i++) {
    int row = 2 * n - 1 - i;
    if (av_left && row < n) { ref[i] = p.at(x0 - 1, y0 + row); avail[i] = true; }
  }
  if (av_above && av_left) { ref[corner] = p.at(x0 - 1, y0 - 1); avail[corner] = true; }
  for (int j = 0; j < 2 * n; j++)
    if (av_above && (j < n || av_ar)) { ref[corner + 1 + j] = p.at(x0 + j, y0 - 1); avail[corner + 1 + j] = true; }
  int len = 4 * n + 1, first = -1;
  for (int i = 0; i < len; i++) if (avail[i]) { first = i; break; }
  if (first < 0) {
    for (int i = 0; i < len; i++) ref[i] = kDCPred;
  } else {
    for (int i = 0; i < first; i++) ref[i] = ref[first];
    for (int i = first + 1; i < len; i++) if (!avail[i]) ref[i] = ref[i - 1];
  }
  auto top = [&](int x) { return ref[corner + 1 + x]; };
  auto left = [&](int y) { return ref[corner - 1 - y]; };
  switch (mode) {
    case kModeV:
      for (int r = 0; r < n; r++) for (int c = 0; c < n; c++) pred[r * n + c] = top(c);
      return;
    case kModeH:
      for (int r = 0; r < n; r++) for (int c = 0; c < n; c++) pred[r * n + c] = left(r);
      return;
    case kModeTM:
      for (int r = 0; r < n; r++) for (int c = 0; c < n; c++) pred[r * n + c] = clamp_byte(top(c) + left(r) - ref[corner]);
      return;
    case kModeDC: {
      int32_t sum = 0;
      for (int i = 0; i < n; i++) sum += top(i) + left(i);
      int32_t dc = (sum + n) / (2 * n);
      for (int i = 0; i < n * n; i++) pred[i] = dc;
      return;
    }
    default: break;
  }
  const AngMode m = kAngModes[mode];
  int32_t ext[3 * kMB + 2] = {0};
  int off = n;
  for (int k = 0; k <= 2 * n; k++) ext[off + k] = m.vertical ? ref[corner + k] : ref[corner - k];
  if (m.angle < 0) {
    int low = (n * m.angle) >> 5;
    for (int x = -1; x >= low; x--) {
      int idx = (x * m.inv_angle + 128) >> 8;
      ext[off + x] = m.vertical ? ref[corner - idx] : ref[corner + idx];
    }
  }
  for (int a = 0; a < n; a++) {
    int pos = (a + 1) * m.angle, idx = pos >> 5, fact = pos & 31;
    for (int b = 0; b < n; b++) {
      int32_t v = ((32 - fact) * ext[off + b + idx + 1] + fact * ext[off + b + idx + 2] + 16) >> 5;
      if (m.vertical) pred[a * n + b] = v; else pred[b * n + a] = v;
    }
  }
}

// ---- motion compensation (motioncomp.go), clamped generic form ----------------------
inline int clampc(int v, int hi) { return v < 0 ? 0 : v >= hi ? hi - 1 : v; }
inline int32_t g(const Plane& r, int x, int y) { return r.d[static_cast<size_t>(clampc(y, r.h)) * r.stride + clampc(x, r.w)]; }
int32_t luma_subpel(const Plane& r, int px, int py, int fx, int fy) {
  if (fx == 0 && fy == 0) return g(r, px, py);
  if (fy == 0) {
    const int32_t* t = kLumaTaps[fx];
    return clamp_byte((t[0] * g(r, px - 1, py) + t[1] * g(r, px, py) + t[2] * g(r, px + 1, py) + t[3] * g(r, px + 2, py) + 32) >> 6);
  }
  if (fx == 0) {
    const int32_t* t = kLumaTaps[fy];
    return clamp_byte((t[0] * g(r, px, py - 1) + t[1] * g(r, px, py) + t[2] * g(r, px, py + 1) + t[3] * g(r, px, py + 2) + 32) >> 6);
  }
  const int32_t *tx = kLumaTaps[fx], *ty = kLumaTaps[fy];
  int32_t h[4];
  for (int k = 0; k < 4; k++) {
    int yy = py - 1 + k;
    h[k] = tx[0] * g(r, px - 1, yy) + tx[1] * g(r, px, yy) + tx[2] * g(r, px + 1, yy) + tx[3] * g(r, px + 2, yy);
  }
  return clamp_byte((ty[0] * h[0] + ty[1] * h[1] + ty[2] * h[2] + ty[3] * h[3] + (1 << 11)) >> 12);
}
int32_t chroma_subpel(const Plane& r, int px, int py, int fx, int fy) {
  if (fx == 0 && fy == 0) return g(r, px, py);
  int32_t v = (8 - fx) * (8 - fy) * g(r, px, py) + fx * (8 - fy) * g(r, px + 1, py) + (8 - fx) * fy * g(r, px, py + 1) +
              fx * fy * g(r, px + 1, py + 1);
  return (v + 32) >> 6;
}
// mcLumaRect / mcChromaRect into a macroblock-local buffer.
void mc_luma_rect(const Plane& r, int mbx, int mby, int px, int py, int pw, int ph, MV mv, int32_t* out) {
  int sx = mbx * kMB + px + (mv.x >> 2), sy = mby * kMB + py + (mv.y >> 2);
  int fx = mv.x & 3, fy = mv.y & 3;
  if (sx - 1 >= 0 && sy - 1 >= 0 && sx + pw + 2 <= r.w && sy + ph + 2 <= r.h) {
    // Interior: every tap is in range, so no coordinate clamp is needed.
    const int s = r.stride;
    const int32_t *tx = kLumaTaps[fx], *ty = kLumaTaps[fy];
    for (int y = 0; y < ph; y++) {
      const uint8_t* row = &r.d[static_cast<size_t>(sy + y) * s + sx];
      int32_t* o = out + (py + y) * kMB + px;
      if (fx == 0 && fy == 0) {
        for (int x = 0; x < pw; x++) o[x] = row[x];
      } else if (fy == 0) {
        for (int x = 0; x < pw; x++)
          o[x] = clamp_byte((tx[0] * row[x - 1] + tx[1] * row[x] + tx[2] * row[x + 1] + tx[3] * row[x + 2] + 32) >> 6);
      } else if (fx == 0) {
        for (int x = 0; x < pw; x++)
          o[x] = clamp_byte((ty[0] * row[x - s] + ty[1] * row[x] + ty[2] * row[x + s] + ty[3] * row[x + 2 * s] + 32) >> 6);
      } else {
        for (int x = 0; x < pw; x++) {
          int32_t h[4];
          for (int k = 0; k < 4; k++) {
            const uint8_t* b = row + (k - 1) * s + x;
            h[k] = tx[0] * b[-1] + tx[1] * b[0] + tx[2] * b[1] + tx[3] * b[2];
          }
          o[x] = clamp_byte((ty[0] * h[0] + ty[1] * h[1] + ty[2] * h[2] + ty[3] * h[3] + (1 << 11)) >> 12);
        }
      }
    }
    return;
  }
  for (int y = 0; y < ph; y++)
    for (int x = 0; x < pw; x++) out[(py + y) * kMB + px + x] = luma_subpel(r, sx + x, sy + y, fx, fy);
}
void mc_chroma_rect(const Plane& r, int cx, int cy, int px, int py, int pw, int ph, MV mv, int32_t* out) {
  int cpx = px / 2, cpy = py / 2;
  int sx = cx + cpx + (mv.x >> 3), sy = cy + cpy + (mv.y >> 3);
  for (int y = 0; y < ph / 2; y++)
    for (int x = 0; x < pw / 2; x++) out[(cpy + y) * kCMB + cpx + x] = chroma_subpel(r, sx + x, sy + y, mv.x & 7, mv.y & 7);
}

// ---- MV prediction (mv.go:21-65) ------------------------------------------------
struct TileBounds { int x0, x1, y0, y1; };
int16_t median3(int16_t a, int16_t b, int16_t c) {
  if (a > b) std::swap(a, b);
  if (b > c) b = c;
  if (a > b) b = a;
  return b;
}
MV predict_mv(const std::vector<MV>& mv, int stride, int mbx, int mby, TileBounds b) {
  MV a, bv, c;
  bool aok = mbx > b.x0, bok = mby > b.y0, cok = false;
  if (aok) a = mv[mby * stride + mbx - 1];
  if (bok) {
    bv = mv[(mby - 1) * stride + mbx];
    if (mbx + 1 < b.x1) { c = mv[(mby - 1) * stride + mbx + 1]; cok = true; }
    else if (aok) { c = mv[(mby - 1) * stride + mbx - 1]; cok = true; }
  }
  if (!aok && !bok && !cok) return MV{};
  if (aok && !bok && !cok) return a;
  if (!aok && bok && !cok) return bv;
  return MV{median3(a.x, bv.x, c.x), median3(a.y, bv.y, c.y)};
}

// ---- rANS encoder and tables (rans.go) ---------------------------------------------
std::vector<uint32_t> normalize_freqs(const std::vector<uint32_t>& counts) {  // rans.go:46-85
  uint64_t total = 0;
  for (uint32_t c : counts) total += c;
  if (total == 0) return {};
  std::vector<uint32_t> f(counts.size(), 0);
  uint32_t used = 0;
  size_t largest = 0;
  for (size_t s = 0; s < counts.size(); s++) {
    if (counts[s] == 0) continue;
    uint32_t v = static_cast<uint32_t>(static_cast<uint64_t>(counts[s]) * kRansM / total);
    if (v == 0) v = 1;
    f[s] = v;
    used += v;
    if (f[s] > f[largest]) largest = s;
  }
  while (used > kRansM) {
    int best = -1;
    for (size_t s = 0; s < f.size(); s++)
      if (f[s] > 1 && (best < 0 || f[s] > f[best])) best = static_cast<int>(s);
    if (best < 0)
      for (size_t s = 0; s < f.size(); s++) if (f[s] > 0) { best = static_cast<int>(s); break; }
    f[best]--;
    used--;
  }
  if (used < kRansM) f[largest] += kRansM - used;
  return f;
}

struct EncTable { uint32_t start[kMaxAlphabet], freq[kMaxAlphabet]; };

std::vector<uint8_t> rans_encode(const std::vector<Tok>& toks, const std::vector<EncTable>& tables) {  // rans.go:199-214
  std::vector<uint8_t> st;
  uint32_t state[kRansStates];
  for (auto& x : state) x = kRansLower;
  for (size_t jj = toks.size(); jj-- > 0;) {
    const Tok& t = toks[jj];
    uint32_t freq = tables[t.ctx].freq[t.sym], start = tables[t.ctx].start[t.sym];
    uint32_t& x = state[jj % kRansStates];
    uint32_t xmax = ((kRansLower >> kRansScaleBits) << 8) * freq;
    while (x >= xmax) { st.push_back(static_cast<uint8_t>(x)); x >>= 8; }
    x = ((x / freq) << kRansScaleBits) + (x % freq) + start;
  }
  for (int i = 0; i < kRansStates; i++)
    for (int k = 0; k < 4; k++) st.push_back(static_cast<uint8_t>(state[i] >> (8 * k)));
  return st;
}

using Freqs = std::vector<std::vector<uint32_t>>;  // empty vector = nil

std::vector<uint8_t> serialize_tables(const Freqs& freqs, const Freqs* prev) {  // rans.go:339-377
  std::vector<uint8_t> out;
  uint8_t tmp[16];
  for (int c = 0; c < kNumContexts; c++) {
    const auto& f = freqs[c];
    if (c == kCtxBypass || f.empty()) { out.push_back(0); continue; }
    const std::vector<uint32_t>* pf = prev ? &(*prev)[c] : nullptr;
    if (pf && pf->empty()) pf = nullptr;
    if (pf && *pf == f) { out.push_back(1); continue; }
    std::vector<uint8_t> full;
    for (uint32_t v : f) { size_t n = put_uvarint(tmp, v); full.insert(full.end(), tmp, tmp + n); }
    if (pf) {
      std::vector<uint8_t> delta;
      for (size_t s = 0; s < f.size(); s++) {
        size_t n = put_varint(tmp, static_cast<int64_t>(f[s]) - static_cast<int64_t>((*pf)[s]));
        delta.insert(delta.end(), tmp, tmp + n);
      }
      if (delta.size() < full.size()) { out.push_back(3); out.insert(out.end(), delta.begin(), delta.end()); continue; }
    }
    out.push_back(2);
    out.insert(out.end(), full.begin(), full.end());
  }
  return out;
}

void append_uvarint(std::vector<uint8_t>& out, uint64_t v) {
  uint8_t tmp[16];
  size_t n = put_uvarint(tmp, v);
  out.insert(out.end(), tmp, tmp + n);
}

// ---- frame encoder (encode_intra.go, encode_inter.go) -----------------------------------
struct SearchCache {
  MV last[4][4];
  MV golden;
  bool try_subs = false, try_golden = false;
};

struct FrameEncoder;

struct TileEncoder {
  FrameEncoder* fe;
  int32_t pred[256], residual[256], coeffs[256], levels[256], recon[256];
  TileStream luma_toks, cb_toks, cr_toks;
  explicit TileEncoder(FrameEncoder* f) : fe(f) {}

  int rd_bits(const std::vector<Tok>& toks) const;
  int64_t code_residual(TileStream& s, int plane, int tx, const Plane& src, Plane& rec, int x0, int y0, int n);
  int64_t pred_err(const Plane& src, const Plane& rec, int x0, int y0, int n, int mode, bool ab, bool lf, bool ar);
  int choose_intra_mode(const Plane& src, const Plane& rec, int x0, int y0, int n, bool ab, bool lf, bool ar);
  int64_t code_luma_mb(TileStream& s, int mbx, int mby, TileBounds b, int tx);
  int64_t luma_mb_cost(int mbx, int mby, TileBounds b, int tx);
  void encode_luma_mb(TileStream& s, int mbx, int mby, TileBounds b);
  void encode_chroma_mb(TileStream& s, int mbx, int mby, TileBounds b);
  int64_t code_luma_mb_best(TileStream& s, int mbx, int mby, TileBounds b);
  int64_t encode_chroma_mb_to(TileStream& s, int mbx, int mby, TileBounds b);
  void encode_intra_tile(TileStream& s, TileBounds b);
  // inter
  int64_t rect_sad(const FrameBuf& ref, int mbx, int mby, int px, int py, int pw, int ph, MV mv);
  MV search_mv_rect(const FrameBuf& ref, int mbx, int mby, int px, int py, int pw, int ph, MV pred);
  int64_t mc_mb_sse(const FrameBuf& ref, int mbx, int mby, MV mv);
  void recon_mc(int mbx, int mby, MV mv);
  int64_t code_inter_luma(TileStream& s, int mbx, int mby, const int32_t* mc, int tx);
  int64_t code_inter_mb(TileStream& s, const FrameBuf& ref, int mbx, int mby, int part, const MV* mvs, MV mvpred);
  void encode_inter_mb(TileStream& s, int mbx, int mby, TileBounds b);
  void encode_inter_tile(TileStream& s, TileBounds b);
};

struct FrameEncoder {
  int W, H;
  FrameBuf src;
  FrameBufP rec;
  FrameBufP ref, golden;  // null for intra frames (golden may be null for the override)
  Quant q;
  int grid_cols, grid_rows;
  std::vector<int> col_edges, row_edges;
  std::vector<TileStream> streams;
  int64_t lambda;
  double lambda_f;
  const CostModel* rdoq = nullptr;
  std::vector<MV> mv;
  std::vector<SearchCache> search;
  int threads;

  FrameEncoder(const SrcImage& img, FrameBufP ref_, FrameBufP golden_, int qp, int tcols, int trows, int nthreads)
      : W(img.W), H(img.H), src(img.W, img.H), rec(std::make_shared<FrameBuf>(img.W, img.H)), ref(std::move(ref_)),
        golden(std::move(golden_)), q(qp), threads(nthreads) {
    load_plane(src.y, img.y, img.W, img.H);
    load_plane(src.cb, img.cb, img.CW, img.CH);
    load_plane(src.cr, img.cr, img.CW, img.CH);
    // newTileGrid, frame.go:87-111
    grid_cols = std::min(std::max(tcols, 1), src.mbc);
    grid_rows = std::min(std::max(trows, 1), src.mbr);
    for (int i = 0; i <= grid_cols; i++) col_edges.push_back(src.mbc * i / grid_cols);
    for (int i = 0; i <= grid_rows; i++) row_edges.push_back(src.mbr * i / grid_rows);
    int64_t qac = q.ac;
    lambda = qac * qac / 6;
    lambda_f = static_cast<double>(lambda);
    mv.assign(static_cast<size_t>(src.mbc) * src.mbr, MV{});
    if (ref) search.assign(mv.size(), SearchCache{});
  }
  int tiles() const { return grid_cols * grid_rows; }
  TileBounds bounds(int i) const {
    int tc = i % grid_cols, tr = i / grid_cols;
    return {col_edges[tc], col_edges[tc + 1], row_edges[tr], row_edges[tr + 1]};
  }
  void encode_tiles(bool inter) {
    streams.assign(tiles(), TileStream{});
    std::atomic<int> next{0};
    auto work = [&]() {
      TileEncoder te(this);
      for (;;) {
        int i = next.fetch_add(1);
        if (i >= tiles()) break;
        if (inter) te.encode_inter_tile(streams[i], bounds(i));
        else te.encode_intra_tile(streams[i], bounds(i));
      }
    };
    int nt = std::min(threads, tiles());
    if (nt <= 1) { work(); return; }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) th.emplace_back(work);
    for (auto& t : th) t.join();
  }
  // measureCosts, encode_intra.go:200-227
  CostModel measure_costs() const {
    std::vector<std::vector<uint32_t>> counts(kNumContexts);
    for (int c = 0; c < kNumContexts; c++) counts[c].assign(alphabet_size(c), 0);
    for (const auto& s : streams) for (const auto& t : s.toks) counts[t.ctx][t.sym]++;
    CostModel cost(kNumContexts);
    for (int c = 0; c < kNumContexts; c++) {
      auto f = normalize_freqs(counts[c]);
      for (int s = 0; s < kMaxAlphabet; s++) cost[c][s] = 0;
      for (int s = 0; s < alphabet_size(c); s++) {
        uint32_t v = f.empty() ? 0 : f[s];
        cost[c][s] = v == 0 ? 20.0 : go_log2(static_cast<double>(kRansM) / static_cast<double>(v));
      }
    }
    return cost;
  }
  void encode(bool inter, bool two_pass, CostModel& cost_store) {  // encode_intra.go:187-195
    encode_tiles(inter);
    if (!two_pass) return;
    cost_store = measure_costs();
    rdoq = &cost_store;
    rec = std::make_shared<FrameBuf>(W, H);  // resetForPass2
    std::fill(mv.begin(), mv.end(), MV{});
    encode_tiles(inter);
  }
  // finish, encode_intra.go:144-168
  std::vector<uint8_t> finish(int qp, const Freqs* prev, Freqs& freqs_out) {
    std::vector<std::vector<uint32_t>> counts(kNumContexts);
    for (int c = 0; c < kNumContexts; c++) counts[c].assign(alphabet_size(c), 0);
    for (const auto& s : streams) for (const auto& t : s.toks) counts[t.ctx][t.sym]++;
    freqs_out.assign(kNumContexts, {});
    std::vector<EncTable> tables(kNumContexts);
    for (int c = 0; c < kNumContexts; c++) {
      EncTable& t = tables[c];
      std::memset(&t, 0, sizeof t);
      if (c == kCtxBypass) continue;
      freqs_out[c] = normalize_freqs(counts[c]);
      uint32_t cum = 0;
      for (size_t s = 0; s < freqs_out[c].size(); s++) { t.start[s] = cum; t.freq[s] = freqs_out[c][s]; cum += freqs_out[c][s]; }
    }
    tables[kCtxBypass].start[0] = 0; tables[kCtxBypass].freq[0] = kRansM / 2;
    tables[kCtxBypass].start[1] = kRansM / 2; tables[kCtxBypass].freq[1] = kRansM / 2;
    std::vector<uint8_t> tb = serialize_tables(freqs_out, prev);
    std::vector<uint8_t> out;
    out.push_back(static_cast<uint8_t>(qp));  // codedHeader.appendTo, frame.go:188-195
    append_uvarint(out, grid_cols);
    append_uvarint(out, grid_rows);
    append_uvarint(out, tb.size());  // assemblePayload, frame.go:224-236
    out.insert(out.end(), tb.begin(), tb.end());
    for (const auto& s : streams) {
      auto bytes = rans_encode(s.toks, tables);
      append_uvarint(out, bytes.size());
      out.insert(out.end(), bytes.begin(), bytes.end());
    }
    return out;
  }
};

int TileEncoder::rd_bits(const std::vector<Tok>& toks) const {  // encode_intra.go:407-446
  if (fe->rdoq) {
    double b = 0;
    for (const auto& t : toks) b += (*fe->rdoq)[t.ctx][t.sym];
    return static_cast<int>(b + 0.5);
  }
  int bits = 0;
  for (const auto& t : toks) {
    if (t.ctx >= kCtxTokenBase) bits += t.sym == kTEOB ? 2 : t.sym == kTZero ? 1 : t.sym == kTEscape ? 6 : 4;
    else if (t.ctx == kCtxSign || t.ctx == kCtxBypass) bits += 1;
    else if (t.ctx == kCtxMVClassX || t.ctx == kCtxMVClassY || t.ctx == kCtxEscClass) bits += 3;
    else if (t.ctx == kCtxMBMode || t.ctx == kCtxCBP) bits += 2;
    else bits += 3;
  }
  return bits;
}

int64_t TileEncoder::code_residual(TileStream& s, int plane, int tx, const Plane& src, Plane& rec, int x0, int y0, int n) {
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) residual[r * n + c] = static_cast<int32_t>(src.at(x0 + c, y0 + r)) - pred[r * n + c];
  forward_dct(residual, coeffs, n);
  if (fe->rdoq) quantize_rdoq(fe->q, coeffs, levels, n, plane, tx, fe->lambda_f, *fe->rdoq);
  else fe->q.quantize(coeffs, levels, n);
  apply_sdh(levels, n);
  encode_block(s, plane, tx, levels, n);
  fe->q.dequantize(levels, coeffs, n);
  inverse_dct(coeffs, recon, n);
  int64_t sse = 0;
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) {
      int32_t rv = clamp_byte(pred[r * n + c] + recon[r * n + c]);
      rec.set(x0 + c, y0 + r, static_cast<uint8_t>(rv));
      int64_t d = static_cast<int64_t>(rv) - src.at(x0 + c, y0 + r);
      sse += d * d;
    }
  return sse;
}

int64_t TileEncoder::pred_err(const Plane& src, const Plane& rec, int x0, int y0, int n, int mode, bool ab, bool lf, bool ar) {
  predict_intra(rec, x0, y0, n, mode, ab, lf, ar, pred);
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) residual[r * n + c] = static_cast<int32_t>(src.at(x0 + c, y0 + r)) - pred[r * n + c];
  return satd(residual, n);
}

int TileEncoder::choose_intra_mode(const Plane& src, const Plane& rec, int x0, int y0, int n, bool ab, bool lf, bool ar) {
  int best = kModeDC;
  int64_t best_err = int64_t(1) << 62;
  for (int m = 0; m < kNumIntraModes; m++) {
    int64_t e = pred_err(src, rec, x0, y0, n, m, ab, lf, ar);
    if (e < best_err) { best_err = e; best = m; }
  }
  return best;
}

int64_t TileEncoder::code_luma_mb(TileStream& s, int mbx, int mby, TileBounds b, int tx) {  // encode_intra.go:288-309
  int n = 4 << tx;
  int tl = b.x0 * kMB, tt = b.y0 * kMB, tr = b.x1 * kMB, mbt = mby * kMB, mbr = (mbx + 1) * kMB;
  int64_t sse = 0;
  for (int by = 0; by < kMB; by += n)
    for (int bx = 0; bx < kMB; bx += n) {
      int x0 = mbx * kMB + bx, y0 = mby * kMB + by;
      bool ab = y0 > tt, lf = x0 > tl, ar = above_right_avail(x0, y0, n, tt, tr, mbt, mbr);
      int mode = choose_intra_mode(fe->src.y, fe->rec->y, x0, y0, n, ab, lf, ar);
      s.put(kCtxLumaMode, mode);
      predict_intra(fe->rec->y, x0, y0, n, mode, ab, lf, ar, pred);
      sse += code_residual(s, 0, tx, fe->src.y, fe->rec->y, x0, y0, n);
    }
  return sse;
}

int64_t TileEncoder::luma_mb_cost(int mbx, int mby, TileBounds b, int tx) {
  TileStream trial;
  int64_t sse = code_luma_mb(trial, mbx, mby, b, tx);
  return sse + fe->lambda * rd_bits(trial.toks);
}

void TileEncoder::encode_luma_mb(TileStream& s, int mbx, int mby, TileBounds b) {  // encode_intra.go:264-274
  int best_tx = 0;
  int64_t best = int64_t(1) << 62;
  for (int tx = 0; tx < 3; tx++) {
    int64_t c = luma_mb_cost(mbx, mby, b, tx);
    if (c < best) { best = c; best_tx = tx; }
  }
  s.put(kCtxTxSize, best_tx);
  code_luma_mb(s, mbx, mby, b, best_tx);
}

void TileEncoder::encode_chroma_mb(TileStream& s, int mbx, int mby, TileBounds b) {  // encode_intra.go:311-334
  int tl = b.x0 * kCMB, tt = b.y0 * kCMB, x0 = mbx * kCMB, y0 = mby * kCMB;
  bool ab = y0 > tt, lf = x0 > tl;
  int best = kModeDC;
  int64_t best_err = int64_t(1) << 62;
  for (int m = 0; m < kNumIntraModes; m++) {
    int64_t e = pred_err(fe->src.cb, fe->rec->cb, x0, y0, kCMB, m, ab, lf, false) +
                pred_err(fe->src.cr, fe->rec->cr, x0, y0, kCMB, m, ab, lf, false);
    if (e < best_err) { best_err = e; best = m; }
  }
  s.put(kCtxChromaMode, best);
  predict_intra(fe->rec->cb, x0, y0, kCMB, best, ab, lf, false, pred);
  code_residual(s, 1, 0, fe->src.cb, fe->rec->cb, x0, y0, kCMB);
  predict_intra(fe->rec->cr, x0, y0, kCMB, best, ab, lf, false, pred);
  code_residual(s, 1, 0, fe->src.cr, fe->rec->cr, x0, y0, kCMB);
}

void TileEncoder::encode_intra_tile(TileStream& s, TileBounds b) {
  for (int mby = b.y0; mby < b.y1; mby++)
    for (int mbx = b.x0; mbx < b.x1; mbx++) {
      encode_luma_mb(s, mbx, mby, b);
      encode_chroma_mb(s, mbx, mby, b);
    }
}

int64_t TileEncoder::code_luma_mb_best(TileStream& s, int mbx, int mby, TileBounds b) {  // encode_inter.go:167-176
  int best_tx = 0;
  int64_t best = int64_t(1) << 62;
  for (int tx = 0; tx < 3; tx++) {
    int64_t c = luma_mb_cost(mbx, mby, b, tx);
    if (c < best) { best = c; best_tx = tx; }
  }
  s.put(kCtxTxSize, best_tx);
  return code_luma_mb(s, mbx, mby, b, best_tx);
}

int64_t plane_sse(const Plane& a, const Plane& b, int x0, int y0, int n) {
  int64_t s = 0;
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) {
      int64_t d = static_cast<int64_t>(a.at(x0 + c, y0 + r)) - b.at(x0 + c, y0 + r);
      s += d * d;
    }
  return s;
}

int64_t TileEncoder::encode_chroma_mb_to(TileStream& s, int mbx, int mby, TileBounds b) {
  encode_chroma_mb(s, mbx, mby, b);
  int x0 = mbx * kCMB, y0 = mby * kCMB;
  return plane_sse(fe->src.cb, fe->rec->cb, x0, y0, kCMB) + plane_sse(fe->src.cr, fe->rec->cr, x0, y0, kCMB);
}

int64_t TileEncoder::rect_sad(const FrameBuf& ref, int mbx, int mby, int px, int py, int pw, int ph, MV mv) {  // encode_inter.go:421-454
  int x0 = mbx * kMB + px, y0 = mby * kMB + py;
  if ((mv.x & 3) == 0 && (mv.y & 3) == 0) {
    int sx = x0 + (mv.x >> 2), sy = y0 + (mv.y >> 2);
    if (sx >= 0 && sy >= 0 && sx + pw <= ref.y.w && sy + ph <= ref.y.h) {
      int64_t sad = 0;
      for (int r = 0; r < ph; r++) {
        const uint8_t* a = &fe->src.y.d[static_cast<size_t>(y0 + r) * fe->src.y.stride + x0];
        const uint8_t* bb = &ref.y.d[static_cast<size_t>(sy + r) * ref.y.stride + sx];
        for (int c = 0; c < pw; c++) sad += std::abs(static_cast<int>(a[c]) - static_cast<int>(bb[c]));
      }
      return sad;
    }
  }
  int32_t mc[256];
  mc_luma_rect(ref.y, mbx, mby, px, py, pw, ph, mv, mc);
  int64_t sad = 0;
  for (int r = 0; r < ph; r++)
    for (int c = 0; c < pw; c++) sad += abs32(static_cast<int32_t>(fe->src.y.at(x0 + c, y0 + r)) - mc[(py + r) * kMB + px + c]);
  return sad;
}

MV TileEncoder::search_mv_rect(const FrameBuf& ref, int mbx, int mby, int px, int py, int pw, int ph, MV pred) {  // encode_inter.go:381-419
  MV best{};
  int64_t best_sad = rect_sad(ref, mbx, mby, px, py, pw, ph, best);
  auto consider = [&](MV m) {
    int64_t sad = rect_sad(ref, mbx, mby, px, py, pw, ph, m);
    if (sad < best_sad) { best_sad = sad; best = m; return true; }
    return false;
  };
  consider(pred);
  for (int step = 16; step >= 1; step >>= 1) {
    for (;;) {
      int cx = best.x, cy = best.y;
      bool moved = consider(MV{clamp_mv(cx + step * 4), clamp_mv(cy)}) || consider(MV{clamp_mv(cx - step * 4), clamp_mv(cy)}) ||
                   consider(MV{clamp_mv(cx), clamp_mv(cy + step * 4)}) || consider(MV{clamp_mv(cx), clamp_mv(cy - step * 4)});
      if (!moved) break;
    }
  }
  for (int sub : {2, 1}) {
    int bx = best.x, by = best.y;
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) consider(MV{clamp_mv(bx + dx * sub), clamp_mv(by + dy * sub)});
  }
  return best;
}

int64_t TileEncoder::mc_mb_sse(const FrameBuf& ref, int mbx, int mby, MV mv) {  // encode_inter.go:202-231
  int32_t mcy[256], mcc[64];
  mc_luma_rect(ref.y, mbx, mby, 0, 0, kMB, kMB, mv, mcy);
  int64_t sse = 0;
  int x0 = mbx * kMB, y0 = mby * kMB;
  for (int r = 0; r < kMB; r++)
    for (int c = 0; c < kMB; c++) { int64_t d = static_cast<int64_t>(fe->src.y.at(x0 + c, y0 + r)) - mcy[r * kMB + c]; sse += d * d; }
  int cx = mbx * kCMB, cy = mby * kCMB;
  mc_chroma_rect(ref.cb, cx, cy, 0, 0, kMB, kMB, mv, mcc);
  for (int r = 0; r < kCMB; r++)
    for (int c = 0; c < kCMB; c++) { int64_t d = static_cast<int64_t>(fe->src.cb.at(cx + c, cy + r)) - mcc[r * kCMB + c]; sse += d * d; }
  mc_chroma_rect(ref.cr, cx, cy, 0, 0, kMB, kMB, mv, mcc);
  for (int r = 0; r < kCMB; r++)
    for (int c = 0; c < kCMB; c++) { int64_t d = static_cast<int64_t>(fe->src.cr.at(cx + c, cy + r)) - mcc[r * kCMB + c]; sse += d * d; }
  return sse;
}

void TileEncoder::recon_mc(int mbx, int mby, MV mv) {  // encode_inter.go:235-260
  int32_t mcy[256], mcc[64];
  const FrameBuf& ref = *fe->ref;
  FrameBuf& rec = *fe->rec;
  mc_luma_rect(ref.y, mbx, mby, 0, 0, kMB, kMB, mv, mcy);
  for (int r = 0; r < kMB; r++) for (int c = 0; c < kMB; c++) rec.y.set(mbx * kMB + c, mby * kMB + r, clamp_byte(mcy[r * kMB + c]));
  int cx = mbx * kCMB, cy = mby * kCMB;
  mc_chroma_rect(ref.cb, cx, cy, 0, 0, kMB, kMB, mv, mcc);
  for (int r = 0; r < kCMB; r++) for (int c = 0; c < kCMB; c++) rec.cb.set(cx + c, cy + r, clamp_byte(mcc[r * kCMB + c]));
  mc_chroma_rect(ref.cr, cx, cy, 0, 0, kMB, kMB, mv, mcc);
  for (int r = 0; r < kCMB; r++) for (int c = 0; c < kCMB; c++) rec.cr.set(cx + c, cy + r, clamp_byte(mcc[r * kCMB + c]));
}

int64_t TileEncoder::code_inter_luma(TileStream& s, int mbx, int mby, const int32_t* mc, int tx) {  // encode_inter.go:359-374
  int n = 4 << tx;
  int64_t sse = 0;
  for (int by = 0; by < kMB; by += n)
    for (int bx = 0; bx < kMB; bx += n) {
      for (int r = 0; r < n; r++) for (int c = 0; c < n; c++) pred[r * n + c] = mc[(by + r) * kMB + bx + c];
      sse += code_residual(s, 0, tx, fe->src.y, fe->rec->y, mbx * kMB + bx, mby * kMB + by, n);
    }
  return sse;
}

bool all_zero_blocks(const TileStream& s) {
  for (const auto& t : s.toks) if (t.sym != kTEOB || t.ctx < kCtxTokenBase) return false;
  return true;
}

int64_t TileEncoder::code_inter_mb(TileStream& s, const FrameBuf& ref, int mbx, int mby, int part, const MV* mvs, MV mvpred) {  // encode_inter.go:267-342
  s.put(kCtxPart, part);
  for (int i = 0; i < kPartCount[part]; i++) {
    MV pp = i > 0 ? mvs[0] : mvpred;
    put_class_value(s, kCtxMVClassX, zigzag_enc(static_cast<int>(mvs[i].x) - pp.x));
    put_class_value(s, kCtxMVClassY, zigzag_enc(static_cast<int>(mvs[i].y) - pp.y));
  }
  int32_t mcy[256], mccb[64], mccr[64];
  int cx = mbx * kCMB, cy = mby * kCMB;
  for (int i = 0; i < kPartCount[part]; i++) {
    const int* rc = kPartRect[part][i];
    mc_luma_rect(ref.y, mbx, mby, rc[0], rc[1], rc[2], rc[3], mvs[i], mcy);
  }
  for (int i = 0; i < kPartCount[part]; i++) {
    const int* rc = kPartRect[part][i];
    mc_chroma_rect(ref.cb, cx, cy, rc[0], rc[1], rc[2], rc[3], mvs[i], mccb);
    mc_chroma_rect(ref.cr, cx, cy, rc[0], rc[1], rc[2], rc[3], mvs[i], mccr);
  }
  int best_tx = 0;
  int64_t best_cost = int64_t(1) << 62;
  for (int tx = 0; tx < 3; tx++) {
    TileStream trial;
    int64_t sse = code_inter_luma(trial, mbx, mby, mcy, tx);
    int64_t cost = sse + fe->lambda * rd_bits(trial.toks);
    if (cost < best_cost) { best_cost = cost; best_tx = tx; }
  }
  luma_toks.toks.clear();
  int64_t sse = code_inter_luma(luma_toks, mbx, mby, mcy, best_tx);
  bool luma_zero = all_zero_blocks(luma_toks);
  cb_toks.toks.clear();
  std::memcpy(pred, mccb, sizeof mccb);
  sse += code_residual(cb_toks, 1, 0, fe->src.cb, fe->rec->cb, cx, cy, kCMB);
  bool cb_zero = all_zero_blocks(cb_toks);
  cr_toks.toks.clear();
  std::memcpy(pred, mccr, sizeof mccr);
  sse += code_residual(cr_toks, 1, 0, fe->src.cr, fe->rec->cr, cx, cy, kCMB);
  bool cr_zero = all_zero_blocks(cr_toks);
  int cbp = (luma_zero ? 0 : kCBPLuma) | (cb_zero ? 0 : kCBPCb) | (cr_zero ? 0 : kCBPCr);
  s.put(kCtxCBP, cbp);
  if (!luma_zero) { s.put(kCtxTxSize, best_tx); s.toks.insert(s.toks.end(), luma_toks.toks.begin(), luma_toks.toks.end()); }
  if (!cb_zero) s.toks.insert(s.toks.end(), cb_toks.toks.begin(), cb_toks.toks.end());
  if (!cr_zero) s.toks.insert(s.toks.end(), cr_toks.toks.begin(), cr_toks.toks.end());
  return sse;
}

void TileEncoder::encode_inter_mb(TileStream& s, int mbx, int mby, TileBounds b) {  // encode_inter.go:37-163
  int idx = mby * fe->src.mbc + mbx;
  MV pred = predict_mv(fe->mv, fe->src.mbc, mbx, mby, b);
  int64_t skip_sse = mc_mb_sse(*fe->ref, mbx, mby, pred);
  int64_t skip_cost = skip_sse + fe->lambda * 2;
  if (skip_sse <= fe->lambda * 8) {
    s.put(kCtxMBMode, kMBSkip);
    recon_mc(mbx, mby, pred);
    fe->mv[idx] = pred;
    return;
  }
  SearchCache& cache = fe->search[idx];
  bool pass2 = fe->rdoq != nullptr;
  int best_part = kPartWhole;
  MV best_mvs[4];
  if (!pass2) cache.last[kPartWhole][0] = search_mv_rect(*fe->ref, mbx, mby, 0, 0, kMB, kMB, pred);
  best_mvs[0] = cache.last[kPartWhole][0];
  int64_t best_inter;
  {
    TileStream trial;
    int64_t sse = code_inter_mb(trial, *fe->ref, mbx, mby, kPartWhole, best_mvs, pred);
    best_inter = sse + fe->lambda * (rd_bits(trial.toks) + 3);
  }
  if (!pass2) {
    int64_t whole_sad = rect_sad(*fe->ref, mbx, mby, 0, 0, kMB, kMB, best_mvs[0]);
    cache.try_subs = whole_sad > 2 * kMB * kMB;
    cache.try_golden = fe->golden && rect_sad(*fe->golden, mbx, mby, 0, 0, kMB, kMB, MV{}) < whole_sad;
    if (cache.try_golden) cache.golden = search_mv_rect(*fe->golden, mbx, mby, 0, 0, kMB, kMB, MV{});
  }
  if (cache.try_subs) {
    for (int part = kPartWhole + 1; part < 4; part++) {
      MV mvs[4];
      for (int i = 0; i < kPartCount[part]; i++) {
        if (pass2) { mvs[i] = cache.last[part][i]; continue; }
        const int* rc = kPartRect[part][i];
        MV pp = i > 0 ? mvs[0] : pred;
        mvs[i] = search_mv_rect(*fe->ref, mbx, mby, rc[0], rc[1], rc[2], rc[3], pp);
        cache.last[part][i] = mvs[i];
      }
      TileStream trial;
      int64_t sse = code_inter_mb(trial, *fe->ref, mbx, mby, part, mvs, pred);
      int64_t cost = sse + fe->lambda * (rd_bits(trial.toks) + 3);
      if (cost < best_inter) { best_inter = cost; best_part = part; std::copy(mvs, mvs + 4, best_mvs); }
    }
  }
  int64_t golden_cost = int64_t(1) << 62;
  MV golden_mv{};
  if (cache.try_golden) {
    MV gmv = cache.golden;
    TileStream trial;
    int64_t sse = code_inter_mb(trial, *fe->golden, mbx, mby, kPartWhole, &gmv, MV{});
    golden_cost = sse + fe->lambda * (rd_bits(trial.toks) + 3);
    golden_mv = gmv;
  }
  int64_t intra_cost = int64_t(1) << 62;
  if (best_inter > fe->lambda * 32) {
    TileStream intra_tok;
    int64_t isse = code_luma_mb_best(intra_tok, mbx, mby, b) + encode_chroma_mb_to(intra_tok, mbx, mby, b);
    intra_cost = isse + fe->lambda * (rd_bits(intra_tok.toks) + 2);
  }
  if (skip_cost <= best_inter && skip_cost <= intra_cost && skip_cost <= golden_cost) {
    s.put(kCtxMBMode, kMBSkip);
    recon_mc(mbx, mby, pred);
    fe->mv[idx] = pred;
  } else if (golden_cost < best_inter && golden_cost <= intra_cost) {
    s.put(kCtxMBMode, kMBInter);
    s.put(kCtxRef, kRefGolden);
    code_inter_mb(s, *fe->golden, mbx, mby, kPartWhole, &golden_mv, MV{});
    fe->mv[idx] = MV{};
  } else if (best_inter <= intra_cost) {
    s.put(kCtxMBMode, kMBInter);
    s.put(kCtxRef, kRefLast);
    code_inter_mb(s, *fe->ref, mbx, mby, best_part, best_mvs, pred);
    fe->mv[idx] = best_mvs[0];
  } else {
    s.put(kCtxMBMode, kMBIntra);
    encode_luma_mb(s, mbx, mby, b);
    encode_chroma_mb(s, mbx, mby, b);
    fe->mv[idx] = MV{};
  }
}

void TileEncoder::encode_inter_tile(TileStream& s, TileBounds b) {
  for (int mby = b.y0; mby < b.y1; mby++)
    for (int mbx = b.x0; mbx < b.x1; mbx++) encode_inter_mb(s, mbx, mby, b);
}

// ---- container writer (writer.go) ------------------------------------------------------
uint32_t crc32_ieee(const uint8_t* p, size_t n) {
  static uint32_t tab[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = c & 1 ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      tab[i] = c;
    }
    init = true;
  }
  uint32_t c = 0xFFFFFFFFu;
  for (size_t i = 0; i < n; i++) c = tab[(c ^ p[i]) & 0xff] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}
void put_le(std::vector<uint8_t>& o, uint64_t v, int bytes) {
  for (int i = 0; i < bytes; i++) o.push_back(static_cast<uint8_t>(v >> (8 * i)));
}

struct IndexEntry { uint64_t off; uint32_t size; bool key; };

// One coded payload's token streams as the rANS coder consumed them, with the
// final frequency vectors (for the grammar walk in the tests).
struct PayloadLog {
  int packet;  // index of the record the payload went into
  bool inter;
  int grid_cols, grid_rows, mbc, mbr;
  std::vector<std::vector<Tok>> tiles;
  Freqs freqs;
};

// ---- encoder ------------------------------------------------------------------------------
struct Encoder {
  int W, H, CW, CH;
  uint32_t fps_num, fps_den;
  int tile_cols, tile_rows, qp, gop, altref, coding;  // coding: 0 raw, 1 coded
  bool fast;
  int threads;
  bool keep_recon;
  std::vector<uint8_t> out;
  std::vector<IndexEntry> index;
  FrameBufP ref, golden;
  Freqs prev;  // last shipped vector per context
  std::vector<uint8_t> prev_y;
  std::vector<SrcImage> pending;
  std::vector<std::vector<uint8_t>> recons;  // packed visible planes per written frame
  std::vector<std::vector<uint8_t>> packets;  // frame payloads in stream order
  CostModel cost_store;
  int64_t accepted = 0;
  bool log_tokens = false;
  std::vector<PayloadLog> logs;

  void write_record(const std::vector<uint8_t>& payload, bool key) {  // writer.go:56-88
    index.push_back({out.size(), static_cast<uint32_t>(payload.size()), key});
    put_le(out, 0xEB17, 2);
    put_le(out, key ? 1 : 0, 2);
    put_le(out, payload.size(), 4);
    out.insert(out.end(), payload.begin(), payload.end());
    packets.push_back(payload);
  }
  std::vector<uint8_t> frame_header(bool key, int cod, bool gov) const {  // format.go:359-373
    std::vector<uint8_t> h;
    h.push_back(static_cast<uint8_t>((key ? 0 : 1) | (cod & 7) << 2 | (gov ? 1 << 5 : 0)));
    if (key) { put_le(h, W, 2); put_le(h, H, 2); h.push_back(0); }
    return h;
  }
  void store_recon(const FrameBuf& f) {
    if (!keep_recon) return;
    std::vector<uint8_t> r(static_cast<size_t>(W) * H + 2 * static_cast<size_t>(CW) * CH);
    for (int y = 0; y < H; y++) std::memcpy(&r[static_cast<size_t>(y) * W], &f.y.d[static_cast<size_t>(y) * f.y.stride], W);
    size_t o = static_cast<size_t>(W) * H;
    for (int y = 0; y < CH; y++) std::memcpy(&r[o + static_cast<size_t>(y) * CW], &f.cb.d[static_cast<size_t>(y) * f.cb.stride], CW);
    o += static_cast<size_t>(CW) * CH;
    for (int y = 0; y < CH; y++) std::memcpy(&r[o + static_cast<size_t>(y) * CW], &f.cr.d[static_cast<size_t>(y) * f.cr.stride], CW);
    recons.push_back(std::move(r));
  }
  void merge(const Freqs& f) {  // mergeFreqs, encoder.go:310-316
    for (int c = 0; c < kNumContexts; c++) if (!f[c].empty()) prev[c] = f[c];
  }
  bool scene_cut(const SrcImage& img) {  // encoder.go:208-227
    bool first = prev_y.empty();
    int64_t sad = 0;
    if (!first)
      for (size_t i = 0; i < img.y.size(); i++) sad += std::abs(static_cast<int>(img.y[i]) - static_cast<int>(prev_y[i]));
    prev_y = img.y;
    return !first && sad > static_cast<int64_t>(W) * H * 28;
  }
  std::unique_ptr<SrcImage> build_altref(const std::vector<SrcImage>& frames) {  // encoder.go:328-407
    int n = std::min<int>(altref, static_cast<int>(frames.size()));
    if (n < 3) return nullptr;
    const SrcImage& t = frames[0];
    std::vector<uint32_t> sy(static_cast<size_t>(W) * H), scb(static_cast<size_t>(CW) * CH), scr(scb.size());
    std::vector<uint16_t> cy(sy.size()), cc(scb.size());
    auto acc = [](std::vector<uint32_t>& sum, std::vector<uint16_t>* cnt, const std::vector<uint8_t>& src, int pw, int x0,
                  int y0, int x1, int y1) {
      for (int y = y0; y < y1; y++)
        for (int x = x0; x < x1; x++) {
          sum[static_cast<size_t>(y) * pw + x] += src[static_cast<size_t>(y) * pw + x];
          if (cnt) (*cnt)[static_cast<size_t>(y) * pw + x]++;
        }
    };
    acc(sy, &cy, t.y, W, 0, 0, W, H);
    acc(scb, &cc, t.cb, CW, 0, 0, CW, CH);
    acc(scr, nullptr, t.cr, CW, 0, 0, CW, CH);
    int accepted_blocks = 0;
    for (int j = 1; j < n; j++) {
      const SrcImage& f = frames[j];
      for (int by = 0; by < H; by += kMB) {
        int bh = std::min(kMB, H - by);
        for (int bx = 0; bx < W; bx += kMB) {
          int bw = std::min(kMB, W - bx);
          int64_t sad = 0;
          for (int y = by; y < by + bh; y++)
            for (int x = bx; x < bx + bw; x++)
              sad += std::abs(static_cast<int>(t.y[static_cast<size_t>(y) * W + x]) - static_cast<int>(f.y[static_cast<size_t>(y) * W + x]));
          if (sad > static_cast<int64_t>(bw) * bh * 6) continue;
          accepted_blocks++;
          acc(sy, &cy, f.y, W, bx, by, bx + bw, by + bh);
          int cx0 = bx / 2, cy0 = by / 2, cx1 = std::min(CW, (bx + bw + 1) / 2), cy1 = std::min(CH, (by + bh + 1) / 2);
          acc(scb, &cc, f.cb, CW, cx0, cy0, cx1, cy1);
          acc(scr, nullptr, f.cr, CW, cx0, cy0, cx1, cy1);
        }
      }
    }
    if (accepted_blocks == 0) return nullptr;
    auto o = std::make_unique<SrcImage>();
    o->W = W; o->H = H; o->CW = CW; o->CH = CH;
    o->y.resize(sy.size()); o->cb.resize(scb.size()); o->cr.resize(scb.size());
    for (size_t i = 0; i < sy.size(); i++) { uint32_t c = cy[i]; o->y[i] = static_cast<uint8_t>((sy[i] + c / 2) / c); }
    for (size_t i = 0; i < scb.size(); i++) {
      uint32_t c = cc[i];
      o->cb[i] = static_cast<uint8_t>((scb[i] + c / 2) / c);
      o->cr[i] = static_cast<uint8_t>((scr[i] + c / 2) / c);
    }
    return o;
  }
  std::vector<uint8_t> encode_frame(const SrcImage& img, FrameBufP r, FrameBufP gld, const Freqs* pv, Freqs& freqs,
                                    FrameBufP& rec_out) {
    FrameEncoder fe(img, r, gld, qp, tile_cols, tile_rows, threads);
    fe.encode(r != nullptr, !fast, cost_store);
    auto payload = fe.finish(qp, pv, freqs);
    rec_out = fe.rec;
    if (log_tokens) {
      PayloadLog l{static_cast<int>(packets.size()), r != nullptr, fe.grid_cols, fe.grid_rows, fe.src.mbc, fe.src.mbr,
                   {}, freqs};
      for (auto& st : fe.streams) l.tiles.push_back(std::move(st.toks));
      logs.push_back(std::move(l));
    }
    return payload;
  }
  void flush_gop() {  // encoder.go:250-306
    if (pending.empty()) return;
    std::vector<SrcImage> frames;
    frames.swap(pending);
    auto alt = build_altref(frames);
    Freqs freqs;
    FrameBufP rec;
    auto payload = encode_frame(frames[0], nullptr, nullptr, nullptr, freqs, rec);
    prev.assign(kNumContexts, {});
    merge(freqs);
    ref = rec;
    golden = rec;
    auto buf = frame_header(true, 1, alt != nullptr);
    if (alt) {
      Freqs ovf;
      FrameBufP ovrec;
      auto ov = encode_frame(*alt, rec, nullptr, &prev, ovf, ovrec);
      merge(ovf);
      golden = ovrec;
      append_uvarint(buf, payload.size());
      buf.insert(buf.end(), payload.begin(), payload.end());
      buf.insert(buf.end(), ov.begin(), ov.end());
    } else {
      buf.insert(buf.end(), payload.begin(), payload.end());
    }
    write_record(buf, true);
    store_recon(*rec);
    for (size_t i = 1; i < frames.size(); i++) {
      Freqs f;
      FrameBufP r;
      auto p = encode_frame(frames[i], ref, golden, &prev, f, r);
      merge(f);
      ref = r;
      auto b = frame_header(false, 2, false);
      b.insert(b.end(), p.begin(), p.end());
      write_record(b, false);
      store_recon(*r);
    }
  }
  void write_frame(SrcImage&& img) {  // encoder.go:158-176, 234-245
    if (coding == 0) {
      auto b = frame_header(true, 0, false);
      b.insert(b.end(), img.y.begin(), img.y.end());
      b.insert(b.end(), img.cb.begin(), img.cb.end());
      b.insert(b.end(), img.cr.begin(), img.cr.end());
      write_record(b, true);
      if (keep_recon) {
        std::vector<uint8_t> r(img.y);
        r.insert(r.end(), img.cb.begin(), img.cb.end());
        r.insert(r.end(), img.cr.begin(), img.cr.end());
        recons.push_back(std::move(r));
      }
      return;
    }
    if (scene_cut(img) && pending.size() >= 8) flush_gop();
    pending.push_back(std::move(img));
    if (static_cast<int>(pending.size()) >= gop) flush_gop();
    accepted++;
  }
  void close() {  // encoder.go:447-452, writer.go:92-119
    flush_gop();
    uint64_t index_off = out.size();
    std::vector<uint8_t> ib;
    for (const auto& e : index) { put_le(ib, e.off, 8); put_le(ib, e.size, 4); put_le(ib, e.key ? 1 : 0, 4); }
    out.insert(out.end(), ib.begin(), ib.end());
    put_le(out, index_off, 8);
    put_le(out, index.size(), 4);
    put_le(out, crc32_ieee(ib.data(), ib.size()), 4);
    out.push_back('E'); out.push_back('B'); out.push_back('I'); out.push_back('X');
  }
};

}  // namespace
}  // namespace ebiv

using ebiv::Encoder;

extern "C" {

// Starts a stream: NewEncoder(w, Config{W,H,fps,tiles}, WithIntra(qp) when
// coding, WithGOP(gop), WithAltRef(altref), WithFastEncode() when fast).
// tile_cols/tile_rows <= 1 both mean a single tile (no auto tiles).
void* ebivenc_new(int w, int h, uint32_t fps_num, uint32_t fps_den, int tile_cols, int tile_rows, int qp, int gop,
                  int altref, int coding, int fast, int threads, int keep_recon) {
  if (w <= 0 || h <= 0 || w > 65535 || h > 65535 || fps_num == 0 || fps_den == 0) return nullptr;
  auto* e = new Encoder();
  e->W = w; e->H = h; e->CW = (w + 1) / 2; e->CH = (h + 1) / 2;
  e->fps_num = fps_num; e->fps_den = fps_den;
  e->tile_cols = tile_cols; e->tile_rows = tile_rows;
  e->qp = qp; e->gop = gop >= 1 ? gop : 1; e->altref = altref >= 0 ? altref : 0;
  e->coding = coding; e->fast = fast != 0; e->threads = threads < 1 ? 1 : threads; e->keep_recon = keep_recon != 0;
  // FileHeader.marshal, format.go:186-204
  auto& o = e->out;
  o.push_back('E'); o.push_back('B'); o.push_back('I'); o.push_back('V');
  ebiv::put_le(o, 4, 2);
  ebiv::put_le(o, 48, 2);
  ebiv::put_le(o, static_cast<uint32_t>(w), 4);
  ebiv::put_le(o, static_cast<uint32_t>(h), 4);
  ebiv::put_le(o, fps_num, 4);
  ebiv::put_le(o, fps_den, 4);
  ebiv::put_le(o, 0, 4);
  o.push_back(0); o.push_back(0); o.push_back(0); o.push_back(0);
  ebiv::put_le(o, static_cast<uint16_t>(tile_cols), 2);
  ebiv::put_le(o, static_cast<uint16_t>(tile_rows), 2);
  ebiv::put_le(o, 0, 4);
  ebiv::put_le(o, 0, 8);
  return e;
}

// Encodes one frame from packed planes (Y w*h, Cb/Cr ceil(w/2)*ceil(h/2)).
int ebivenc_write(void* h, const uint8_t* y, const uint8_t* cb, const uint8_t* cr) {
  auto* e = static_cast<Encoder*>(h);
  ebiv::SrcImage img;
  img.W = e->W; img.H = e->H; img.CW = e->CW; img.CH = e->CH;
  img.y.assign(y, y + static_cast<size_t>(e->W) * e->H);
  img.cb.assign(cb, cb + static_cast<size_t>(e->CW) * e->CH);
  img.cr.assign(cr, cr + static_cast<size_t>(e->CW) * e->CH);
  e->write_frame(std::move(img));
  return 0;
}

// Flushes the GOP and writes index + footer.  Returns the container size.
size_t ebivenc_close(void* h) {
  auto* e = static_cast<Encoder*>(h);
  e->close();
  return e->out.size();
}
const uint8_t* ebivenc_data(void* h) { return static_cast<Encoder*>(h)->out.data(); }
size_t ebivenc_packet_count(void* h) { return static_cast<Encoder*>(h)->packets.size(); }
const uint8_t* ebivenc_packet(void* h, size_t i, size_t* len) {
  auto* e = static_cast<Encoder*>(h);
  *len = e->packets[i].size();
  return e->packets[i].data();
}
size_t ebivenc_recon_count(void* h) { return static_cast<Encoder*>(h)->recons.size(); }
const uint8_t* ebivenc_recon(void* h, size_t i) { return static_cast<Encoder*>(h)->recons[i].data(); }
void ebivenc_free(void* h) { delete static_cast<Encoder*>(h); }

// Token log of the coded payloads (set before the first write): per payload
// its record index, inter flag, tile grid, macroblock grid, the (ctx, sym)
// tokens of each tile in stream order and the frequency vector of every
// context (zeros where the context shipped none).
void ebivenc_log_tokens(void* h, int on) { static_cast<Encoder*>(h)->log_tokens = on != 0; }
size_t ebivenc_payload_count(void* h) { return static_cast<Encoder*>(h)->logs.size(); }
int ebivenc_payload_info(void* h, size_t i, int32_t* info) {
  auto* e = static_cast<Encoder*>(h);
  if (i >= e->logs.size()) return -1;
  const auto& l = e->logs[i];
  info[0] = l.packet; info[1] = l.inter; info[2] = l.grid_cols; info[3] = l.grid_rows; info[4] = l.mbc; info[5] = l.mbr;
  return 0;
}
const uint16_t* ebivenc_payload_tokens(void* h, size_t i, int tile, size_t* count) {
  auto* e = static_cast<Encoder*>(h);
  *count = 0;
  if (i >= e->logs.size() || tile < 0 || tile >= static_cast<int>(e->logs[i].tiles.size())) return nullptr;
  const auto& t = e->logs[i].tiles[tile];
  *count = t.size();
  return reinterpret_cast<const uint16_t*>(t.data());
}
int ebivenc_payload_freqs(void* h, size_t i, uint32_t* out) {  // kNumContexts x kMaxAlphabet
  auto* e = static_cast<Encoder*>(h);
  if (i >= e->logs.size()) return -1;
  const auto& f = e->logs[i].freqs;
  for (int c = 0; c < ebiv::kNumContexts; c++)
    for (int k = 0; k < ebiv::kMaxAlphabet; k++)
      out[c * ebiv::kMaxAlphabet + k] = c < static_cast<int>(f.size()) && k < static_cast<int>(f[c].size()) ? f[c][k] : 0;
  return 0;
}

// Test hooks over the encoder's own block grammar (residual.go:30-75,
// rans.go:199-214) so the decoder can be checked on hand-written level
// patterns: encodes `count` blocks of n*n levels (row-major) after sign data
// hiding (when sdh != 0), returning the serialized key-frame tables and the
// tile stream.  levels is updated in place with the SDH-adjusted values.
int ebivenc_blocks(int plane, int tx, int n, int count, int32_t* levels, int sdh, uint8_t* tables_out,
                   size_t* tables_len, uint8_t* stream_out, size_t* stream_len) {
  ebiv::TileStream s;
  for (int b = 0; b < count; b++) {
    int32_t* lv = levels + static_cast<size_t>(b) * n * n;
    if (sdh) ebiv::apply_sdh(lv, n);
    ebiv::encode_block(s, plane, tx, lv, n);
  }
  ebiv::Freqs freqs(ebiv::kNumContexts);
  std::vector<std::vector<uint32_t>> counts(ebiv::kNumContexts);
  for (int c = 0; c < ebiv::kNumContexts; c++) counts[c].assign(ebiv::alphabet_size(c), 0);
  for (const auto& t : s.toks) counts[t.ctx][t.sym]++;
  std::vector<ebiv::EncTable> tables(ebiv::kNumContexts);
  for (int c = 0; c < ebiv::kNumContexts; c++) {
    std::memset(&tables[c], 0, sizeof tables[c]);
    if (c == ebiv::kCtxBypass) continue;
    freqs[c] = ebiv::normalize_freqs(counts[c]);
    uint32_t cum = 0;
    for (size_t k = 0; k < freqs[c].size(); k++) { tables[c].start[k] = cum; tables[c].freq[k] = freqs[c][k]; cum += freqs[c][k]; }
  }
  tables[ebiv::kCtxBypass].freq[0] = tables[ebiv::kCtxBypass].freq[1] = ebiv::kRansM / 2;
  tables[ebiv::kCtxBypass].start[1] = ebiv::kRansM / 2;
  auto tb = ebiv::serialize_tables(freqs, nullptr);
  auto st = ebiv::rans_encode(s.toks, tables);
  if (tb.size() > *tables_len || st.size() > *stream_len) return -1;
  std::memcpy(tables_out, tb.data(), tb.size());
  std::memcpy(stream_out, st.data(), st.size());
  *tables_len = tb.size();
  *stream_len = st.size();
  return static_cast<int>(s.toks.size());
}

// Random-token round trip (rans_test.go:43-75): encodes the given (ctx, sym)
// stream; returns serialized tables and the tile stream.
int ebivenc_tokens(const int* ctxs, const int* syms, int count, uint8_t* tables_out, size_t* tables_len,
                   uint8_t* stream_out, size_t* stream_len) {
  std::vector<ebiv::Tok> toks(count);
  std::vector<std::vector<uint32_t>> counts(ebiv::kNumContexts);
  for (int c = 0; c < ebiv::kNumContexts; c++) counts[c].assign(ebiv::alphabet_size(c), 0);
  for (int i = 0; i < count; i++) {
    toks[i] = {static_cast<uint16_t>(ctxs[i]), static_cast<uint16_t>(syms[i])};
    counts[ctxs[i]][syms[i]]++;
  }
  ebiv::Freqs freqs(ebiv::kNumContexts);
  std::vector<ebiv::EncTable> tables(ebiv::kNumContexts);
  for (int c = 0; c < ebiv::kNumContexts; c++) {
    std::memset(&tables[c], 0, sizeof tables[c]);
    if (c == ebiv::kCtxBypass) continue;
    freqs[c] = ebiv::normalize_freqs(counts[c]);
    uint32_t cum = 0;
    for (size_t k = 0; k < freqs[c].size(); k++) { tables[c].start[k] = cum; tables[c].freq[k] = freqs[c][k]; cum += freqs[c][k]; }
  }
  tables[ebiv::kCtxBypass].freq[0] = tables[ebiv::kCtxBypass].freq[1] = ebiv::kRansM / 2;
  tables[ebiv::kCtxBypass].start[1] = ebiv::kRansM / 2;
  auto tb = ebiv::serialize_tables(freqs, nullptr);
  auto st = ebiv::rans_encode(toks, tables);
  if (tb.size() > *tables_len || st.size() > *stream_len) return -1;
  std::memcpy(tables_out, tb.data(), tb.size());
  std::memcpy(stream_out, st.data(), st.size());
  *tables_len = tb.size();
  *stream_len = st.size();
  return 0;
}

// forwardDCT / inverseDCT of the encoder (transform.go:44-79) for the
// transform round-trip test.
int ebivenc_forward_dct(const int32_t* src, int32_t* dst, int n) {
  if (n != 4 && n != 8 && n != 16) return -1;
  ebiv::forward_dct(src, dst, n);
  return 0;
}

// quantizer.quantize / dequantize (quant.go:55-80) for the quantiser test.
int ebivenc_quantize(int qp, const int32_t* coeffs, int32_t* levels, int n) {
  ebiv::Quant q(qp);
  q.quantize(coeffs, levels, n);
  return 0;
}

// normalizeFreqs (rans.go:46-85) for the invariant test.
int ebivenc_normalize(const uint32_t* counts, int n, uint32_t* out) {
  std::vector<uint32_t> c(counts, counts + n);
  auto f = ebiv::normalize_freqs(c);
  if (f.empty()) return 0;
  std::memcpy(out, f.data(), sizeof(uint32_t) * n);
  return 1;
}

}  // extern "C"
