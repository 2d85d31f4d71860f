// ebiv_kernels.hip — gfx950 kernels of the EBIV decode path.
//
//   ebiv_parse       one wavefront per tile: static-table interleaved rANS
//                    (ebiv/rans.go:222-310) and the macroblock grammar
//                    (ebiv/decode_intra.go:121-167, ebiv/decode_inter.go:8-62,
//                    ebiv/residual.go:79-126, ebiv/mv.go:21-128).  Emits one
//                    MBRec per macroblock plus a sparse coefficient list.
//   ebiv_recon_inter one wavefront per skip/inter macroblock: dequantisation
//                    (ebiv/quant.go:72-80), inverse DCT (ebiv/transform.go:129-194),
//                    motion compensation (ebiv/motioncomp.go) and the clamped
//                    add (ebiv/predict.go:209-219).
//   ebiv_recon_intra one workgroup per tile: intra macroblocks on an anti-diagonal
//                    wavefront (left, above-left, above, above-right dependencies),
//                    each macroblock's block chain inside one wavefront
//                    (ebiv/predict.go:50-203).
//   ebiv_rgba        YCbCr 4:2:0 -> RGBA (frame.go:8-67).
//
// Every wavefront-cooperative routine runs with all 64 lanes active; the
// decoder state is wave-uniform and lane-level work is explicit.
#include <hip/hip_runtime.h>

#include "ebiv_device.h"
#include "ebiv_format.h"

namespace ebiv {

__constant__ int32_t c_dct[3][256];   // buildDCTMatrix for n = 4, 8, 16
__constant__ uint8_t c_zz[3][256];    // zigzag scan (row-major positions)
__constant__ int32_t c_taps[4][4] = {{0, 64, 0, 0}, {-6, 55, 18, -3}, {-8, 40, 40, -8}, {-3, 18, 55, -6}};
__constant__ int32_t c_part_count[4] = {1, 2, 2, 4};
// angModes (predict.go:36-43): angle, invAngle, vertical; modes 0-3 unused.
__constant__ int32_t c_ang[kNumIntraModes][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {32, 0, 1},
                                                 {13, 0, 1}, {-13, -630, 1}, {32, 0, 0}, {13, 0, 0}, {-13, -630, 0}};

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ int32_t clamp255(int32_t v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
__device__ __forceinline__ int sidx_of(int n) { return n == 4 ? 0 : (n == 8 ? 1 : 2); }
__device__ __forceinline__ uint32_t rd_lane(uint32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }

// ============================================================================
// Entropy kernel
// ============================================================================

// Descending byte reader over a tile stream.  Each lane holds one dword of a
// 256-byte window, so reading a byte is a readlane; the next lower window is
// loaded one window ahead.
struct ByteWin {
  uintptr_t tile;      // absolute address of tile byte 0
  uintptr_t lo, hi;    // dword-aligned readable range [lo, hi)
  uintptr_t wbase;     // 256-aligned base of the current window
  uint32_t cur, nxt;

  __device__ __forceinline__ uint32_t fetch(uintptr_t wb, int lane) const {
    uintptr_t a = wb + 4u * lane;
    return (a >= lo && a < hi) ? *reinterpret_cast<const uint32_t*>(a) : 0u;
  }
  __device__ void init(const uint8_t* t, uint32_t len, int lane) {
    tile = reinterpret_cast<uintptr_t>(t);
    lo = tile & ~uintptr_t(3);
    hi = (tile + len + 3) & ~uintptr_t(3);
    wbase = (tile + (len ? len - 1 : 0)) & ~uintptr_t(255);
    cur = fetch(wbase, lane);
    nxt = fetch(wbase - 256, lane);
  }
  // Byte p of the tile; p only ever decreases between calls.
  __device__ __forceinline__ uint32_t byte(uint32_t p, int lane) {
    uintptr_t a = tile + p;
    if (a < wbase) {
      cur = nxt;
      wbase -= 256;
      nxt = fetch(wbase - 256, lane);
    }
    uint32_t off = static_cast<uint32_t>(a - wbase);
    uint32_t w = rd_lane(cur, static_cast<int>(off >> 2));
    return (w >> ((off & 3u) * 8u)) & 0xFFu;
  }
};

__device__ __forceinline__ int zigzag_dec(uint32_t u) { return static_cast<int>(u >> 1) ^ -static_cast<int>(u & 1u); }
__device__ __forceinline__ int16_t clamp_mv(int v) { return static_cast<int16_t>(v < -kMVClamp ? -kMVClamp : (v > kMVClamp ? kMVClamp : v)); }
__device__ __forceinline__ int dtoken_ctx(int plane, int tx, int band, int prev) {
  return plane == 0 ? kCtxTokenBase + (tx * kNumBands + band) * kNumPrevClass + prev
                    : kCtxTokenBase + kLumaTokenCtx + band * kNumPrevClass + prev;
}
__device__ __forceinline__ int dband(int i) { return i == 0 ? 0 : (i <= 2 ? 1 : (i <= 8 ? 2 : (i <= 20 ? 3 : 4))); }

__device__ __forceinline__ uint32_t coef_entry(int pos, int blk, int32_t level) {
  return static_cast<uint32_t>(pos) | (static_cast<uint32_t>(blk) << 8) | (static_cast<uint32_t>(level) << 13);
}

__device__ __forceinline__ int16_t median3(int16_t a, int16_t b, int16_t c) {
  if (a > b) { int16_t t = a; a = b; b = t; }
  if (b > c) b = c;
  if (a > b) b = a;
  return b;
}

// predictMV (mv.go:21-52) over the LDS rows of the tile's MV grid.
__device__ __forceinline__ uint32_t predict_mv(const uint32_t* prow, const uint32_t* crow, int lx, int tw, bool top_row) {
  bool aok = lx > 0, bok = !top_row, cok = false;
  uint32_t a = 0, b = 0, c = 0;
  if (aok) a = crow[lx - 1];
  if (bok) {
    b = prow[lx];
    if (lx + 1 < tw) { c = prow[lx + 1]; cok = true; }
    else if (aok) { c = prow[lx - 1]; cok = true; }
  }
  if (!aok && !bok && !cok) return 0;
  if (aok && !bok && !cok) return a;
  if (!aok && bok && !cok) return b;
  auto X = [](uint32_t v) { return static_cast<int16_t>(v & 0xFFFFu); };
  auto Y = [](uint32_t v) { return static_cast<int16_t>(v >> 16); };
  int16_t px = median3(X(a), X(b), X(c));
  int16_t py = median3(Y(a), Y(b), Y(c));
  return static_cast<uint32_t>(static_cast<uint16_t>(px)) | (static_cast<uint32_t>(static_cast<uint16_t>(py)) << 16);
}

__device__ __forceinline__ int16_t mv_comp(uint32_t v, int comp) {
  return static_cast<int16_t>(comp ? (v >> 16) : (v & 0xFFFFu));
}
__device__ __forceinline__ uint32_t mv_set(uint32_t v, int comp, int16_t c) {
  uint32_t u = static_cast<uint16_t>(c);
  return comp ? ((v & 0xFFFFu) | (u << 16)) : ((v & 0xFFFF0000u) | u);
}

// Grammar states of the entropy kernel.  Every state up to P_FSIGN decodes
// exactly one symbol in the context it names; the states after it are
// bookkeeping steps that decode nothing.  The order of decodes is that of
// decodeTile / decodeLumaMB / decodeChromaMB (decode_intra.go:121-167),
// decodeInterMB / reconInterMB (decode_inter.go:8-118), getClassValue
// (mv.go:95-108) and decodeBlock (residual.go:79-126).
enum PState : int {
  P_MBMODE, P_REF, P_PART, P_MVCLS, P_MVBITS, P_CBP, P_TX, P_LMODE, P_CMODE,
  P_TOKEN, P_SIGN, P_ESCCLS, P_ESCBITS, P_FSIGN,
  P_ENDBLK,    // a block's tokens are complete: resolve the hidden/explicit first sign
  P_BLKDONE,   // a block is complete: next block, next plane, or end of macroblock
  P_LUMADONE,  // inter luma complete: uncoded chroma is written, coded chroma parsed
  P_ENDMB,     // macroblock complete: write its record, start the next one
  P_DONE
};

__device__ __forceinline__ uint32_t mv_pick(uint32_t m0, uint32_t m1, uint32_t m2, uint32_t m3, int i) {
  return i == 0 ? m0 : (i == 1 ? m1 : (i == 2 ? m2 : m3));
}

__global__ __launch_bounds__(64) void ebiv_parse(const FrameJob* __restrict__ jobs, const WorkItem* __restrict__ items) {
  extern __shared__ uint32_t smem[];
  const WorkItem it = items[blockIdx.x];
  const FrameJob& job = jobs[it.job];
  const uint8_t* const pkt = job.pkt;
  const uint8_t* const desc = job.desc;
  MBRec* const recs = job.recs;
  uint32_t* const coefs = job.coefs;
  TileStatus* const status = job.status;
  const FrameHdr* hdr = reinterpret_cast<const FrameHdr*>(desc);
  const TileDesc* tiles = reinterpret_cast<const TileDesc*>(desc + desc_tiles_off());
  const int lane = lane_id();
  const int t = static_cast<int>(it.idx);
  const TileDesc td = tiles[t];

  // LDS: cumulative tables, then two MV rows of the tile width.
  uint16_t* cum = reinterpret_cast<uint16_t*>(smem);
  const uint32_t* gcum = reinterpret_cast<const uint32_t*>(hdr->cum);
  constexpr int kCumWords = 72 * kCumStride / 2;
  for (int i = lane; i < kCumWords; i += 64) smem[i] = gcum[i];
  uint32_t* prow = smem + kCumWords;
  uint32_t* crow = prow + hdr->max_tile_w;
  const int tw = td.x1 - td.x0;
  __syncthreads();

  if (td.byte_len < 16) {  // ransDecoder.reset (rans.go:237-250)
    if (lane == 0) status[t] = TileStatus{kErrCorrupt | (kTileShort << 8), 0, 0, 0};
    return;
  }
  const uint32_t used0 = hdr->used[0], used1 = hdr->used[1], used2 = hdr->used[2];
  ByteWin win;
  win.init(pkt + td.byte_off, td.byte_len, lane);
  uint32_t pos = td.byte_len;
  uint32_t s0, s1, s2, s3;
  {
    uint32_t sv[4];
#pragma unroll
    for (int i = 3; i >= 0; i--) {  // pop32: big-endian from the top, state 3 first
      uint32_t a = win.byte(--pos, lane);
      uint32_t b = win.byte(--pos, lane);
      uint32_t c = win.byte(--pos, lane);
      uint32_t e = win.byte(--pos, lane);
      sv[i] = (a << 24) | (b << 16) | (c << 8) | e;
    }
    s0 = sv[0]; s1 = sv[1]; s2 = sv[2]; s3 = sv[3];
  }

  // Coefficient output: lane k buffers entry cbase+k until 64 are queued.
  uint32_t* const cdst = coefs + td.coef_base;
  uint32_t cbase = 0, cn = 0, cbuf = 0;

  const int mbc = hdr->mbc;
  const bool inter_frame = hdr->inter != 0;
  int32_t err = 0;
  uint32_t nsym = 0, done = 0, nintra = 0;
  int mbx = td.x0, mby = td.y0;

  // Macroblock state.
  uint32_t kind = 2, part = 0, cbp = 0, tx = 0, refsel = 0, cmode = 0, wmask = 0, ncoef = 0;
  uint32_t lmode0 = 0, lmode1 = 0, coef_start = 0, mvcur = 0, pred = 0;
  uint32_t mv0 = 0, mv1 = 0, mv2 = 0, mv3 = 0;
  int pi = 0, comp = 0, npart = 1;
  uint32_t clsval = 0;
  int cls_bits = 0, cls = 0;
  bool intra_mb = !inter_frame;
  // Block state.
  int plane = 0, blk = 0, n = 4, nb = 16, total = 16, si = 0, prev = 0, first = -1, last = -1;
  int first_pos = 0;
  int32_t first_mag = 0, pend_mag = 0, sum_abs = 0;

  int st = inter_frame ? P_MBMODE : P_TX;
  while (st != P_DONE) {
    // ---------------- bookkeeping states (no symbol) ----------------
    if (st >= P_ENDBLK) {
      if (st == P_ENDBLK) {
        if (first >= 0 && last - first < kSDHMinSpan) {
          st = P_FSIGN;
          continue;
        }
        if (first >= 0) {  // sign hidden in the parity of the level sum
          const uint32_t e = coef_entry(first_pos, blk, (sum_abs & 1) ? -first_mag : first_mag);
          if (static_cast<uint32_t>(lane) == cn) cbuf = e;
          if (++cn == 64) { cdst[cbase + lane] = cbuf; cbase += 64; cn = 0; }
          ncoef++;
        }
        st = P_BLKDONE;
      }
      if (st == P_BLKDONE) {
        wmask |= 1u << blk;
        si = 0; prev = 0; first = -1; last = -1; sum_abs = 0;
        if (plane == 0) {
          if (blk + 1 < nb) {
            blk++;
            st = intra_mb ? P_LMODE : P_TOKEN;
          } else {
            st = intra_mb ? P_CMODE : P_LUMADONE;
          }
        } else if (blk == 16 && (intra_mb || (cbp & kCBPCr))) {
          blk = 17;
          st = P_TOKEN;
        } else {
          st = P_ENDMB;
        }
        if (st != P_LUMADONE && st != P_ENDMB) continue;
      }
      if (st == P_LUMADONE) {
        if (!(cbp & kCBPCb)) wmask |= 1u << 16;
        if (!(cbp & kCBPCr)) wmask |= 1u << 17;
        si = 0; prev = 0; first = -1; last = -1; sum_abs = 0;
        plane = 1; n = 8; total = 64;
        if (cbp & kCBPCb) { blk = 16; st = P_TOKEN; continue; }
        if (cbp & kCBPCr) { blk = 17; st = P_TOKEN; continue; }
        st = P_ENDMB;
      }
      // P_ENDMB: the macroblock record (decode_intra.go:121-135 loop body).
      if (kind == 2) nintra++;
      crow[mbx - td.x0] = mvcur;
      if (lane == 0) {
        MBRec r;
        r.kind = static_cast<uint8_t>(kind);
        r.part = static_cast<uint8_t>(part);
        r.cbp = static_cast<uint8_t>(cbp);
        r.tx = static_cast<uint8_t>(tx);
        r.refsel = static_cast<uint8_t>(refsel);
        r.cmode = static_cast<uint8_t>(cmode);
        r.pad0 = 0;
        r.pad1 = 0;
        r.wmask = wmask;
        r.coef_off = td.coef_base + coef_start;
        r.coef_count = ncoef;
        r.mv[0][0] = mv_comp(mv0, 0); r.mv[0][1] = mv_comp(mv0, 1);
        r.mv[1][0] = mv_comp(mv1, 0); r.mv[1][1] = mv_comp(mv1, 1);
        r.mv[2][0] = mv_comp(mv2, 0); r.mv[2][1] = mv_comp(mv2, 1);
        r.mv[3][0] = mv_comp(mv3, 0); r.mv[3][1] = mv_comp(mv3, 1);
        r.lmode[0] = lmode0;
        r.lmode[1] = lmode1;
        r.pad2 = 0;
        recs[mby * mbc + mbx] = r;
      }
      done++;
      if (err) break;
      if (++mbx == td.x1) {
        mbx = td.x0;
        if (++mby == td.y1) break;
        wave_sync();
        uint32_t* tmp = prow; prow = crow; crow = tmp;
      }
      kind = 2; part = 0; cbp = 0; tx = 0; refsel = 0; cmode = 0; wmask = 0; ncoef = 0;
      lmode0 = 0; lmode1 = 0; mvcur = 0; mv0 = mv1 = mv2 = mv3 = 0;
      coef_start = cbase + cn;
      intra_mb = !inter_frame;
      plane = 0; blk = 0; si = 0; prev = 0; first = -1; last = -1; sum_abs = 0;
      st = inter_frame ? P_MBMODE : P_TX;
      continue;
    }

    // ---------------- one symbol: ransDecoder.decode (rans.go:270-304) ----------------
    int ctx;
    switch (st) {
      case P_MBMODE: ctx = kCtxMBMode; break;
      case P_REF: ctx = kCtxRef; break;
      case P_PART: ctx = kCtxPart; break;
      case P_MVCLS: ctx = comp ? kCtxMVClassY : kCtxMVClassX; break;
      case P_CBP: ctx = kCtxCBP; break;
      case P_TX: ctx = kCtxTxSize; break;
      case P_LMODE: ctx = kCtxLumaMode; break;
      case P_CMODE: ctx = kCtxChromaMode; break;
      case P_TOKEN: ctx = dtoken_ctx(plane, static_cast<int>(tx), dband(si), prev); break;
      case P_SIGN: case P_FSIGN: ctx = kCtxSign; break;
      case P_ESCCLS: ctx = kCtxEscClass; break;
      default: ctx = kCtxBypass; break;  // P_MVBITS, P_ESCBITS
    }
    int v = 0;
    const uint32_t ub = ctx < 32 ? used0 : (ctx < 64 ? used1 : used2);
    if (!((ub >> (ctx & 31)) & 1u)) {
      err = kErrCorrupt | (kTileUnused << 8);
    } else {
      uint32_t x = s0;
      const uint32_t slot = x & (kRansM - 1);
      uint32_t sym, start, freq;
      if (ctx == kCtxBypass) {  // fixed {2048, 2048} (rans.go:317-322)
        sym = slot >> 11;
        start = sym << 11;
        freq = 2048;
      } else {
        // Lane-parallel slot2sym: lane s holds the start of symbol s; the
        // symbol is the last start not above the slot.
        const uint32_t c = cum[ctx * kCumStride + (lane < 16 ? lane : 16)];
        const uint64_t m = __ballot(c <= slot);
        sym = static_cast<uint32_t>(__popcll(m)) - 1u;
        start = rd_lane(c, static_cast<int>(sym));
        freq = rd_lane(c, static_cast<int>(sym) + 1) - start;
      }
      x = freq * (x >> kRansScaleBits) + slot - start;
      while (x < kRansLower) {
        if (pos == 0) {
          err = kErrCorrupt | (kTileUnderrun << 8);
          break;
        }
        --pos;
        x = (x << 8) | win.byte(pos, lane);
      }
      s0 = s1; s1 = s2; s2 = s3; s3 = x;
      nsym++;
      v = static_cast<int>(sym);
    }
    if (err) {
      // The failing macroblock keeps what Go wrote of it: a failed mode
      // decode reads as skip (decode_inter.go:9-16); otherwise only the
      // blocks and planes completed before the failure.
      if (st == P_MBMODE) {
        pred = predict_mv(prow, crow, mbx - td.x0, tw, mby == td.y0);
        kind = 0; mv0 = pred; mvcur = pred; wmask = 0x3FFFFu;
      }
      st = P_ENDMB;
      continue;
    }

    // ---------------- grammar transition ----------------
    bool has_emit = false;
    uint32_t entry = 0;
    int32_t mag = 0;  // > 0: a magnitude was completed this step
    int mvval = -1;   // >= 0: an MV class value was completed this step
    switch (st) {
      case P_MBMODE:
        if (v == kMBSkip) {
          pred = predict_mv(prow, crow, mbx - td.x0, tw, mby == td.y0);
          kind = 0; mv0 = pred; mvcur = pred; wmask = 0x3FFFFu;
          st = P_ENDMB;
        } else if (v == kMBIntra) {
          kind = 2; intra_mb = true; mvcur = 0;
          st = P_TX;
        } else {
          kind = 1;
          st = P_REF;
        }
        break;
      case P_REF:
        refsel = static_cast<uint32_t>(v);
        pred = v == kRefGolden ? 0u : predict_mv(prow, crow, mbx - td.x0, tw, mby == td.y0);
        st = P_PART;
        break;
      case P_PART:
        part = static_cast<uint32_t>(v);
        npart = c_part_count[v];
        pi = 0; comp = 0;
        st = P_MVCLS;
        break;
      case P_MVCLS:
        if (v <= 1) mvval = v;
        else { cls = v; cls_bits = v - 1; clsval = 0; st = P_MVBITS; }
        break;
      case P_MVBITS:
        clsval = (clsval << 1) | static_cast<uint32_t>(v);
        if (--cls_bits == 0) mvval = static_cast<int>((1u << (cls - 1)) | clsval);
        break;
      case P_CBP:
        cbp = static_cast<uint32_t>(v);
        if (cbp & kCBPLuma) {
          st = P_TX;
        } else {
          wmask |= 0xFFFFu;
          st = P_LUMADONE;
        }
        break;
      case P_TX:
        tx = static_cast<uint32_t>(v);
        n = 4 << v; total = n * n; nb = (16 / n) * (16 / n);
        plane = 0; blk = 0; si = 0; prev = 0; first = -1; last = -1; sum_abs = 0;
        st = intra_mb ? P_LMODE : P_TOKEN;
        break;
      case P_LMODE:
        if (blk < 8) lmode0 |= static_cast<uint32_t>(v) << (4 * blk);
        else lmode1 |= static_cast<uint32_t>(v) << (4 * (blk - 8));
        st = P_TOKEN;
        break;
      case P_CMODE:
        cmode = static_cast<uint32_t>(v);
        plane = 1; blk = 16; n = 8; total = 64;
        si = 0; prev = 0; first = -1; last = -1; sum_abs = 0;
        st = P_TOKEN;
        break;
      case P_TOKEN:
        if (v == kTEOB) {
          st = P_ENDBLK;
        } else if (v == kTZero) {
          prev = 0;
          st = (++si == total) ? P_ENDBLK : P_TOKEN;
        } else if (v == kTEscape) {
          st = P_ESCCLS;
        } else {
          prev = v == 1 ? 1 : 2;
          mag = v;
        }
        break;
      case P_SIGN:
        entry = coef_entry(c_zz[sidx_of(n)][si], blk, v ? -pend_mag : pend_mag);
        has_emit = true;
        last = si;
        st = (++si == total) ? P_ENDBLK : P_TOKEN;
        break;
      case P_FSIGN:
        entry = coef_entry(first_pos, blk, v ? -first_mag : first_mag);
        has_emit = true;
        st = P_BLKDONE;
        break;
      case P_ESCCLS:
        if (v <= 1) { prev = 2; mag = 5 + v; }
        else { cls = v; cls_bits = v - 1; clsval = 0; st = P_ESCBITS; }
        break;
      default:  // P_ESCBITS
        clsval = (clsval << 1) | static_cast<uint32_t>(v);
        if (--cls_bits == 0) { prev = 2; mag = 5 + static_cast<int32_t>((1u << (cls - 1)) | clsval); }
        break;
    }
    if (mag > 0) {  // a coefficient magnitude is known (residual.go:99-117)
      sum_abs += mag;
      if (first < 0) {
        first = si;
        first_mag = mag;
        first_pos = c_zz[sidx_of(n)][si];
        last = si;
        st = (++si == total) ? P_ENDBLK : P_TOKEN;
      } else {
        pend_mag = mag;
        st = P_SIGN;
      }
    }
    if (mvval >= 0) {  // an MV delta component is known (decode_inter.go:36-45)
      const int d = zigzag_dec(static_cast<uint32_t>(mvval));
      const uint32_t base = pi > 0 ? mv0 : pred;
      const uint32_t nv = mv_set(mv_pick(mv0, mv1, mv2, mv3, pi), comp, clamp_mv(mv_comp(base, comp) + d));
      if (pi == 0) mv0 = nv; else if (pi == 1) mv1 = nv; else if (pi == 2) mv2 = nv; else mv3 = nv;
      comp ^= 1;
      if (comp == 0 && ++pi == npart) {
        mvcur = refsel == kRefGolden ? 0u : mv0;
        st = P_CBP;
      } else {
        st = P_MVCLS;
      }
    }
    if (has_emit) {
      if (static_cast<uint32_t>(lane) == cn) cbuf = entry;
      if (++cn == 64) { cdst[cbase + lane] = cbuf; cbase += 64; cn = 0; }
      ncoef++;
    }
  }
  if (static_cast<uint32_t>(lane) < cn) cdst[cbase + lane] = cbuf;
  if (lane == 0) status[t] = TileStatus{err, done, nsym, nintra};
}

// ============================================================================
// Reconstruction shared pieces
// ============================================================================

// Per-wave scratch: dense dequantised coefficients, then residuals.
struct WaveScratch {
  int32_t coef[384];
  int32_t tmp[384];
};

// Dequantises the macroblock's coefficient list into dense per-block arrays:
// luma block b of size n at [b*n*n, (b+1)*n*n), Cb at 256, Cr at 320.
__device__ void load_coefs(WaveScratch& ws, const uint32_t* __restrict__ coefs, const MBRec& r, int n, int qdc, int qac,
                           int lane) {
  for (int i = lane; i < 384; i += 64) ws.coef[i] = 0;
  wave_sync();
  uint32_t cnt = r.coef_count < 384u ? r.coef_count : 384u;
  for (uint32_t i = lane; i < cnt; i += 64) {
    uint32_t e = coefs[r.coef_off + i];
    int pos = e & 0xFF, blk = (e >> 8) & 31;
    int32_t level = static_cast<int32_t>(e) >> 13;
    int32_t v = level * (pos == 0 ? qdc : qac);
    int idx;
    if (blk < 16) {
      int nn = n * n;
      idx = blk * nn + (pos < nn ? pos : nn - 1);
      if (idx >= 256) continue;
    } else if (blk < 18) {
      idx = 256 + (blk - 16) * 64 + (pos & 63);
    } else {
      continue;
    }
    ws.coef[idx] = v;
  }
  wave_sync();
}

// Separable inverse DCT of every block (transform.go:129-194): column pass
// into tmp, row pass back into coef, with 64-bit accumulation and the
// per-pass (x + 4096) >> 13 rounding.  Zero coefficients add nothing, so the
// result equals the reference's sparse form bit for bit.
__device__ void idct_all(WaveScratch& ws, int n, bool luma, bool chroma, int lane) {
  const int32_t* cl = c_dct[sidx_of(n)];
  const int32_t* c8 = c_dct[1];
  const int nn = n * n;
  // column pass: tmp[i][l] = sum_k C[k][i] * src[k][l]
  for (int e = lane; e < 384; e += 64) {
    if (e < 256 ? !luma : !chroma) continue;
    int bn, base, local;
    const int32_t* c;
    if (e < 256) { bn = n; base = (e / nn) * nn; local = e - base; c = cl; }
    else { bn = 8; base = 256 + ((e - 256) >> 6) * 64; local = (e - 256) & 63; c = c8; }
    int i = local / bn, l = local - i * bn;
    int64_t acc = 0;
    for (int k = 0; k < bn; k++) acc += static_cast<int64_t>(c[k * bn + i]) * ws.coef[base + k * bn + l];
    ws.tmp[e] = static_cast<int32_t>((acc + 4096) >> 13);
  }
  wave_sync();
  // row pass: out[i][m] = sum_l C[l][m] * tmp[i][l]
  for (int e = lane; e < 384; e += 64) {
    if (e < 256 ? !luma : !chroma) continue;
    int bn, base, local;
    const int32_t* c;
    if (e < 256) { bn = n; base = (e / nn) * nn; local = e - base; c = cl; }
    else { bn = 8; base = 256 + ((e - 256) >> 6) * 64; local = (e - 256) & 63; c = c8; }
    int i = local / bn, mm = local - i * bn;
    int64_t acc = 0;
    for (int l = 0; l < bn; l++) acc += static_cast<int64_t>(c[l * bn + mm]) * ws.tmp[base + i * bn + l];
    ws.coef[e] = static_cast<int32_t>((acc + 4096) >> 13);
  }
  wave_sync();
}

__device__ __forceinline__ int clampc(int v, int hi) { return v < 0 ? 0 : (v >= hi ? hi - 1 : v); }

// Luma quarter-pel sample (motioncomp.go:49-78).  The 1-D and full-pel cases
// are the 2-D filter with a {0,64,0,0} tap row, which reproduces their
// (v+32)>>6 and copy results exactly, and clamped coordinates equal the
// interior fast paths wherever those apply.
__device__ __forceinline__ int32_t luma_mc(const uint8_t* __restrict__ p, int pw, int ph, int x, int y, int fx, int fy) {
  const int32_t* tx = c_taps[fx];
  const int32_t* ty = c_taps[fy];
  int x0 = clampc(x - 1, pw), x1 = clampc(x, pw), x2 = clampc(x + 1, pw), x3 = clampc(x + 2, pw);
  int32_t v = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint8_t* row = p + static_cast<size_t>(clampc(y - 1 + k, ph)) * pw;
    int32_t h = tx[0] * row[x0] + tx[1] * row[x1] + tx[2] * row[x2] + tx[3] * row[x3];
    v += ty[k] * h;
  }
  return clamp255((v + 2048) >> 12);
}

// Chroma eighth-pel bilinear sample (motioncomp.go:82-97).
__device__ __forceinline__ int32_t chroma_mc(const uint8_t* __restrict__ p, int cw, int ch, int x, int y, int fx, int fy) {
  int xa = clampc(x, cw), xb = clampc(x + 1, cw);
  const uint8_t* r0 = p + static_cast<size_t>(clampc(y, ch)) * cw;
  const uint8_t* r1 = p + static_cast<size_t>(clampc(y + 1, ch)) * cw;
  int32_t v = (8 - fx) * (8 - fy) * r0[xa] + fx * (8 - fy) * r0[xb] + (8 - fx) * fy * r1[xa] + fx * fy * r1[xb];
  return (v + 32) >> 6;
}

__device__ __forceinline__ int part_of(int part, int x, int y) {  // partRects, model.go:91-96
  switch (part) {
    case kPart16x8: return y >= 8 ? 1 : 0;
    case kPart8x16: return x >= 8 ? 1 : 0;
    case kPart8x8: return (y >= 8 ? 2 : 0) + (x >= 8 ? 1 : 0);
    default: return 0;
  }
}

__device__ __forceinline__ int tile_of(const uint8_t* desc, const FrameHdr* hdr, int mbx, int mby) {
  const uint16_t* ct = reinterpret_cast<const uint16_t*>(desc + desc_coltile_off(hdr->ntiles));
  const uint16_t* rt = reinterpret_cast<const uint16_t*>(desc + desc_rowtile_off(hdr->ntiles, hdr->mbc));
  return rt[mby] * hdr->grid_cols + ct[mbx];
}

// ============================================================================
// Inter reconstruction: one wavefront per macroblock
// ============================================================================
__global__ __launch_bounds__(256) void ebiv_recon_inter(const FrameJob* __restrict__ jobs,
                                                        const WorkItem* __restrict__ items) {
  __shared__ WaveScratch scratch[4];
  const WorkItem it = items[blockIdx.x];
  const FrameJob& job = jobs[it.job];
  const uint8_t* const desc = job.desc;
  const MBRec* const recs = job.recs;
  const uint32_t* const coefs = job.coefs;
  const TileStatus* const status = job.status;
  const Planes rec = job.rec, ref = job.ref, gold = job.gold;
  const FrameHdr* hdr = reinterpret_cast<const FrameHdr*>(desc);
  const TileDesc* tiles = reinterpret_cast<const TileDesc*>(desc + desc_tiles_off());
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int mbc = hdr->mbc, nmb = mbc * hdr->mbr;
  const int mb = static_cast<int>(it.idx) + wave;
  if (mb >= nmb) return;
  const int mbx = mb % mbc, mby = mb / mbc;
  const int t = tile_of(desc, hdr, mbx, mby);
  const TileDesc td = tiles[t];
  const uint32_t local = static_cast<uint32_t>((mby - td.y0) * (td.x1 - td.x0) + (mbx - td.x0));
  if (local >= status[t].mbs_done) return;
  const MBRec r = recs[mb];
  if (r.kind == 2) return;
  WaveScratch& ws = scratch[wave];
  const Planes& src = (r.kind == 1 && r.refsel == kRefGolden) ? gold : ref;
  const int n = 4 << (r.tx > 2 ? 2 : r.tx);
  const bool has_res = r.kind == 1 && r.coef_count > 0;
  if (has_res) {
    load_coefs(ws, coefs, r, n, hdr->qdc, hdr->qac, lane);
    idct_all(ws, n, (r.cbp & kCBPLuma) != 0, (r.cbp & (kCBPCb | kCBPCr)) != 0, lane);
  }
  const int part = r.kind == 1 ? r.part : 0;
  // Luma: lane -> row lane/4, four pixels from column (lane%4)*4.
  {
    const int y = lane >> 2, x0 = (lane & 3) * 4;
    const int pi = part_of(part, x0, y);
    const int mvx = r.mv[pi][0], mvy = r.mv[pi][1];
    const int fx = mvx & 3, fy = mvy & 3;
    const int sx = mbx * 16 + x0 + (mvx >> 2), sy = mby * 16 + y + (mvy >> 2);
    const int per = 16 / n;
    const int blk = (y / n) * per + x0 / n;
    if ((r.wmask >> blk) & 1u) {
      uint32_t word = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        int32_t p = luma_mc(src.y, src.pw, src.ph, sx + k, sy, fx, fy);
        if (has_res && (r.cbp & kCBPLuma)) {
          int xx = x0 + k;
          p = clamp255(p + ws.coef[blk * n * n + (y % n) * n + (xx % n)]);
        }
        word |= static_cast<uint32_t>(p) << (8 * k);
      }
      *reinterpret_cast<uint32_t*>(rec.y + static_cast<size_t>(mby * 16 + y) * rec.pw + mbx * 16 + x0) = word;
    }
  }
  // Chroma: lanes 0-31 Cb, 32-63 Cr; row (lane%32)/4, two pixels from (lane%4)*2.
  {
    const int pl = lane >> 5, l = lane & 31;
    const int y = l >> 2, x0 = (l & 3) * 2;
    const int pi = part_of(part, x0 * 2, y * 2);
    const int mvx = r.mv[pi][0], mvy = r.mv[pi][1];
    const int fx = mvx & 7, fy = mvy & 7;
    const int sx = mbx * 8 + x0 + (mvx >> 3), sy = mby * 8 + y + (mvy >> 3);
    const uint8_t* sp = pl ? src.cr : src.cb;
    uint8_t* dp = pl ? rec.cr : rec.cb;
    const bool coded = has_res && (r.cbp & (pl ? kCBPCr : kCBPCb));
    if ((r.wmask >> (16 + pl)) & 1u) {
      uint32_t half = 0;
#pragma unroll
      for (int k = 0; k < 2; k++) {
        int32_t p = chroma_mc(sp, src.cw, src.ch, sx + k, sy, fx, fy);
        if (coded) p = clamp255(p + ws.coef[256 + pl * 64 + y * 8 + x0 + k]);
        else p = clamp255(p);
        half |= static_cast<uint32_t>(p) << (8 * k);
      }
      *reinterpret_cast<uint16_t*>(dp + static_cast<size_t>(mby * 8 + y) * rec.cw + mbx * 8 + x0) = static_cast<uint16_t>(half);
    }
  }
}

// ============================================================================
// Intra reconstruction
// ============================================================================

// LDS context of one macroblock: the row above (corner, 16 above, 16
// above-right), the left column, and the macroblock itself as it is rebuilt.
struct IntraScratch {
  WaveScratch ws;
  uint8_t L[17][33];     // [0][*] row above, [r][0] left column, [1..16][1..16] the MB
  uint8_t C[2][9][9];    // chroma: [0][*] corner+above, [r][0] left
  int32_t ref[2][68];    // HEVC reference array per plane in flight
  int32_t ext[2][52];
  int32_t dc[2];
};

// Builds ref[0..4n] for one block (predictIntra + substituteRefs,
// predict.go:63-90, 134-156) with the substitution in closed form: the
// availability pattern is always (below-left never, left, corner, above,
// above-right), so the first available sample and every gap fill are known.
// at(dx, dy) reads the sample at block-relative (dx, dy), dx/dy >= -1.
template <typename At>
__device__ void build_ref_one(int32_t* ref, int n, bool av_above, bool av_left, bool av_ar, At at, int i) {
  if (i > 4 * n) return;
  int32_t v;
  if (!av_left && !av_above) {
    v = kDCPred;
  } else if (i < 2 * n) {
    if (av_left) {
      int r = 2 * n - 1 - i;
      if (r >= n) r = n - 1;
      v = at(-1, r);
    } else {
      v = at(0, -1);
    }
  } else if (i == 2 * n) {
    v = av_left ? (av_above ? at(-1, -1) : at(-1, 0)) : at(0, -1);
  } else {
    int j = i - 2 * n - 1;
    if (av_above) {
      if (j >= n && !av_ar) j = n - 1;
      v = at(j, -1);
    } else {
      v = at(-1, 0);
    }
  }
  ref[i] = v;
}

// All 4n+1 entries with one wavefront (n = 16 needs 65, so lane 0 also does 64).
template <typename At>
__device__ void build_ref(int32_t* ref, int n, bool av_above, bool av_left, bool av_ar, At at, int lane) {
  build_ref_one(ref, n, av_above, av_left, av_ar, at, lane);
  build_ref_one(ref, n, av_above, av_left, av_ar, at, lane + 64);
}

// Prediction of pixel (r, c) of an n x n block from its reference array
// (predict.go:95-128, 162-203).  ext must already hold the angular line.
__device__ __forceinline__ int32_t intra_pixel(const int32_t* ref, const int32_t* ext, int32_t dc, int n, int mode, int r, int c) {
  const int corner = 2 * n;
  switch (mode) {
    case kModeV: return ref[corner + 1 + c];
    case kModeH: return ref[corner - 1 - r];
    case kModeTM: return clamp255(ref[corner + 1 + c] + ref[corner - 1 - r] - ref[corner]);
    case kModeDC: return dc;
    default: break;
  }
  const bool vert = c_ang[mode][2] != 0;
  int a = vert ? r : c;
  int b = vert ? c : r;
  int pos = (a + 1) * c_ang[mode][0], idx = pos >> 5, fact = pos & 31;
  int32_t s0 = ext[n + b + idx + 1], s1 = ext[n + b + idx + 2];
  return ((32 - fact) * s0 + fact * s1 + 16) >> 5;
}

// Angular main line (predict.go:162-186) for the modes that need it, and the
// DC value.  Lanes [0, 64) cooperate; call after ref is complete.
__device__ void prep_angular(const int32_t* ref, int32_t* ext, int32_t* dcout, int n, int mode, int lane) {
  const int corner = 2 * n, off = n;
  if (mode >= kModeD45) {
    const int angle = c_ang[mode][0], inv = c_ang[mode][1];
    const bool vert = c_ang[mode][2] != 0;
    if (lane <= 2 * n) ext[off + lane] = vert ? ref[corner + lane] : ref[corner - lane];
    if (lane == 0) ext[3 * n + 1] = 0;
    if (angle < 0) {
      int low = (n * angle) >> 5;
      int x = -1 - lane;
      if (x >= low) {
        int idx = (x * inv + 128) >> 8;
        ext[off + x] = vert ? ref[corner - idx] : ref[corner + idx];
      }
    }
  } else if (mode == kModeDC && lane == 0) {
    int32_t sum = 0;
    for (int i = 0; i < n; i++) sum += ref[corner + 1 + i] + ref[corner - 1 - i];
    *dcout = (sum + n) / (2 * n);
  }
}

__device__ void intra_mb(IntraScratch& S, const uint32_t* __restrict__ coefs, const MBRec& r, const FrameHdr* hdr,
                         const TileDesc& td, Planes rec, int mbx, int mby, int lane) {
  const int n = 4 << (r.tx > 2 ? 2 : r.tx);
  WaveScratch& ws = S.ws;
  if (r.coef_count > 0) {
    load_coefs(ws, coefs, r, n, hdr->qdc, hdr->qac, lane);
    idct_all(ws, n, true, true, lane);
  } else {
    for (int i = lane; i < 384; i += 64) ws.coef[i] = 0;
  }
  // Load the context: luma row above (x from -1 to 31) and left column, chroma likewise.
  const int X = mbx * 16, Y = mby * 16;
  for (int i = lane; i < 33; i += 64) {
    int x = X - 1 + i, y = Y - 1;
    S.L[0][i] = (x >= 0 && x < rec.pw && y >= 0) ? rec.y[static_cast<size_t>(y) * rec.pw + x] : 0;
  }
  if (lane < 16) {
    int x = X - 1, y = Y + lane;
    S.L[1 + lane][0] = (x >= 0) ? rec.y[static_cast<size_t>(y) * rec.pw + x] : 0;
  }
  {
    const int CX = mbx * 8, CY = mby * 8;
    int pl = lane >> 5, l = lane & 31;
    const uint8_t* cp = pl ? rec.cr : rec.cb;
    if (l < 9) {
      int x = CX - 1 + l, y = CY - 1;
      S.C[pl][0][l] = (x >= 0 && x < rec.cw && y >= 0) ? cp[static_cast<size_t>(y) * rec.cw + x] : 0;
    } else if (l < 17) {
      int rr = l - 9, x = CX - 1, y = CY + rr;
      S.C[pl][1 + rr][0] = (x >= 0) ? cp[static_cast<size_t>(y) * rec.cw + x] : 0;
    }
  }
  wave_sync();

  // Luma blocks in decode order (decode_intra.go:137-157, availability from tile bounds).
  const int tl = td.x0 * 16, tt = td.y0 * 16, tr = td.x1 * 16;
  const int per = 16 / n, nb = per * per;
  for (int b = 0; b < nb; b++) {
    if (!((r.wmask >> b) & 1u)) break;
    const int bx = (b % per) * n, by = (b / per) * n;
    const int x0 = X + bx, y0 = Y + by;
    const int mode = (b < 8 ? (r.lmode[0] >> (4 * b)) : (r.lmode[1] >> (4 * (b - 8)))) & 15;
    const bool av_above = y0 > tt, av_left = x0 > tl;
    bool av_ar;
    if (y0 <= tt || x0 + 2 * n > tr) av_ar = false;  // aboveRightAvailable, predict.go:50-58
    else if (y0 == Y) av_ar = true;
    else av_ar = x0 + 2 * n <= X + 16;
    auto at = [&](int dx, int dy) -> int32_t { return S.L[by + dy + 1][bx + dx + 1]; };
    build_ref(S.ref[0], n, av_above, av_left, av_ar, at, lane);
    wave_sync();
    prep_angular(S.ref[0], S.ext[0], &S.dc[0], n, mode, lane);
    wave_sync();
    for (int p = lane; p < n * n; p += 64) {
      int rr = p / n, cc = p - rr * n;
      int32_t pred = intra_pixel(S.ref[0], S.ext[0], S.dc[0], n, mode, rr, cc);
      int32_t v = clamp255(pred + ws.coef[b * n * n + p]);
      S.L[by + rr + 1][bx + cc + 1] = static_cast<uint8_t>(v);
    }
    wave_sync();
  }
  // Chroma: Cb and Cr share one mode; lanes 0-31 build Cb's reference, 32-63 Cr's.
  const int cmode = r.cmode;
  {
    const int CX = mbx * 8, CY = mby * 8;
    const bool av_above = CY > td.y0 * 8, av_left = CX > td.x0 * 8;
    for (int pl = 0; pl < 2; pl++) {
      auto at = [&](int dx, int dy) -> int32_t { return S.C[pl][dy + 1][dx + 1]; };
      build_ref(S.ref[pl], 8, av_above, av_left, false, at, lane);
    }
    wave_sync();
    const int pl = lane >> 5, l = lane & 31;
    prep_angular(S.ref[pl], S.ext[pl], &S.dc[pl], 8, cmode, l);
    wave_sync();
  }
  // Write back: luma 16 rows x 16, lane -> row lane/4, 4 pixels.
  {
    const int y = lane >> 2, x0 = (lane & 3) * 4;
    const int blk = (y / n) * per + x0 / n;
    if ((r.wmask >> blk) & 1u) {
      uint32_t word = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) word |= static_cast<uint32_t>(S.L[y + 1][x0 + k + 1]) << (8 * k);
      *reinterpret_cast<uint32_t*>(rec.y + static_cast<size_t>(Y + y) * rec.pw + X + x0) = word;
    }
  }
  // Chroma pixels: lane -> (plane, row, 2 pixels) with prediction + residual.
  {
    const int pl = lane >> 5, l = lane & 31;
    const int y = l >> 2, x0 = (l & 3) * 2;
    if ((r.wmask >> (16 + pl)) & 1u) {
      uint32_t half = 0;
#pragma unroll
      for (int k = 0; k < 2; k++) {
        int32_t pred = intra_pixel(S.ref[pl], S.ext[pl], S.dc[pl], 8, cmode, y, x0 + k);
        int32_t v = clamp255(pred + ws.coef[256 + pl * 64 + y * 8 + x0 + k]);
        half |= static_cast<uint32_t>(v) << (8 * k);
      }
      uint8_t* dp = pl ? rec.cr : rec.cb;
      *reinterpret_cast<uint16_t*>(dp + static_cast<size_t>(mby * 8 + y) * rec.cw + mbx * 8 + x0) = static_cast<uint16_t>(half);
    }
  }
  wave_sync();
}

// One workgroup per tile.  Macroblock (x, y) depends on (x-1, y), (x-1, y-1),
// (x, y-1) and (x+1, y-1), so step s = x + 2y (tile-relative) may run all its
// macroblocks at once.
__global__ __launch_bounds__(256) void ebiv_recon_intra(const FrameJob* __restrict__ jobs,
                                                        const WorkItem* __restrict__ items) {
  __shared__ IntraScratch scratch[4];
  const WorkItem it = items[blockIdx.x];
  const FrameJob& job = jobs[it.job];
  const uint8_t* const desc = job.desc;
  const MBRec* const recs = job.recs;
  const uint32_t* const coefs = job.coefs;
  const TileStatus* const status = job.status;
  const Planes rec = job.rec;
  const FrameHdr* hdr = reinterpret_cast<const FrameHdr*>(desc);
  const TileDesc* tiles = reinterpret_cast<const TileDesc*>(desc + desc_tiles_off());
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int t = static_cast<int>(it.idx);
  const TileDesc td = tiles[t];
  const TileStatus st = status[t];
  if (st.n_intra == 0) return;
  const int tw = td.x1 - td.x0, th = td.y1 - td.y0;
  const int steps = (tw - 1) + 2 * (th - 1) + 1;
  const int mbc = hdr->mbc;
  for (int s = 0; s < steps; s++) {
    // Macroblocks on this step: rows ry with rx = s - 2*ry in [0, tw).
    int ry_lo = s - (tw - 1) > 0 ? (s - (tw - 1) + 1) / 2 : 0;
    int ry_hi = s / 2 < th - 1 ? s / 2 : th - 1;
    for (int ry = ry_lo + wave; ry <= ry_hi; ry += 4) {
      int rx = s - 2 * ry;
      uint32_t local = static_cast<uint32_t>(ry * tw + rx);
      if (local >= st.mbs_done) continue;
      int mbx = td.x0 + rx, mby = td.y0 + ry;
      const MBRec r = recs[mby * mbc + mbx];
      if (r.kind != 2) continue;
      intra_mb(scratch[wave], coefs, r, hdr, td, rec, mbx, mby, lane);
    }
    // Macroblocks of the next step read what other wavefronts of this
    // workgroup stored in this one.  The barrier orders the stores, but the
    // CU's vector L1 may still hold lines fetched before them (another
    // wavefront's neighbour loads touch the same 128-byte lines), so the
    // acquire at agent scope invalidates L1 before the next step's loads.
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
}

// ============================================================================
// YCbCr 4:2:0 -> RGBA (frame.go:8-67): four pixels per thread.
// ============================================================================
__device__ __forceinline__ uint32_t rgba_px(int32_t yv, int32_t cbv, int32_t crv) {
  int32_t yy = yv * 0x10101, b1 = cbv - 128, r1 = crv - 128;
  int32_t r = yy + 91881 * r1, g = yy - 22554 * b1 - 46802 * r1, b = yy + 116130 * b1;
  r = r < 0 ? 0 : (r > 0xff0000 ? 0xff0000 : r);
  g = g < 0 ? 0 : (g > 0xff0000 ? 0xff0000 : g);
  b = b < 0 ? 0 : (b > 0xff0000 ? 0xff0000 : b);
  return static_cast<uint32_t>(r >> 16) | (static_cast<uint32_t>(g >> 16) << 8) | (static_cast<uint32_t>(b >> 16) << 16) |
         0xFF000000u;
}

__global__ __launch_bounds__(256) void ebiv_rgba(const uint8_t* __restrict__ y, int ys, const uint8_t* __restrict__ cb,
                                                 const uint8_t* __restrict__ cr, int cs, int w, int h,
                                                 uint32_t* __restrict__ out) {
  const int qw = (w + 3) >> 2;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= qw * h) return;
  const int row = idx / qw, q = idx - row * qw;
  const int col = q * 4;
  const uint8_t* yr = y + static_cast<size_t>(row) * ys;
  const uint8_t* br = cb + static_cast<size_t>(row >> 1) * cs;
  const uint8_t* rr = cr + static_cast<size_t>(row >> 1) * cs;
  uint32_t* o = out + static_cast<size_t>(row) * w + col;
  if (col + 4 <= w) {
    uint4 v;
    v.x = rgba_px(yr[col], br[col >> 1], rr[col >> 1]);
    v.y = rgba_px(yr[col + 1], br[col >> 1], rr[col >> 1]);
    v.z = rgba_px(yr[col + 2], br[(col + 2) >> 1], rr[(col + 2) >> 1]);
    v.w = rgba_px(yr[col + 3], br[(col + 2) >> 1], rr[(col + 2) >> 1]);
    if ((reinterpret_cast<uintptr_t>(o) & 15) == 0) {
      *reinterpret_cast<uint4*>(o) = v;
    } else {
      o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    }
  } else {
    for (int k = 0; col + k < w; k++) o[k] = rgba_px(yr[col + k], br[(col + k) >> 1], rr[(col + k) >> 1]);
  }
}

}  // namespace ebiv

// Host-side launch shims (C linkage inside the library only).
namespace ebiv {

hipError_t upload_tables(const int32_t dct[3][256], const uint8_t zz[3][256]) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_dct), dct, sizeof(int32_t) * 3 * 256);
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_zz), zz, 3 * 256);
}

void launch_parse(hipStream_t s, int nitems, int max_tile_w, const FrameJob* jobs, const WorkItem* items) {
  size_t lds = (72 * kCumStride / 2 + 2 * static_cast<size_t>(max_tile_w)) * sizeof(uint32_t);
  hipLaunchKernelGGL(ebiv_parse, dim3(nitems), dim3(64), lds, s, jobs, items);
}

void launch_recon_inter(hipStream_t s, int nitems, const FrameJob* jobs, const WorkItem* items) {
  hipLaunchKernelGGL(ebiv_recon_inter, dim3(nitems), dim3(256), 0, s, jobs, items);
}

void launch_recon_intra(hipStream_t s, int nitems, const FrameJob* jobs, const WorkItem* items) {
  hipLaunchKernelGGL(ebiv_recon_intra, dim3(nitems), dim3(256), 0, s, jobs, items);
}

void launch_rgba(hipStream_t s, const uint8_t* y, int ys, const uint8_t* cb, const uint8_t* cr, int cs, int w, int h,
                 uint32_t* out) {
  int n = ((w + 3) / 4) * h;
  hipLaunchKernelGGL(ebiv_rgba, dim3((n + 255) / 256), dim3(256), 0, s, y, ys, cb, cr, cs, w, h, out);
}

}  // namespace ebiv
