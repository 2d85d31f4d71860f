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
// partCount (model.go:91-96) as nibbles: 16x16 -> 1, 16x8 -> 2, 8x16 -> 2, 8x8 -> 4.
constexpr uint32_t kPartCountNibbles = 0x4221u;
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
// Uniform values for "s" operands of inline asm: the compiler may hold a
// uniform value in a VGPR, and an "s" operand is not converted for it.
__device__ __forceinline__ uint32_t sgpr(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// (The builtin returns int: each half goes back through uint32_t so the low
// half is zero-extended, not sign-extended, into the 64-bit address.)
template <typename T>
__device__ __forceinline__ T* sgpr(T* p) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  const uint64_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a)));
  const uint64_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a >> 32)));
  return reinterpret_cast<T*>(lo | (hi << 32));
}

// ============================================================================
// Entropy kernel
// ============================================================================

// Descending byte reader over a tile stream.  Each lane holds one dword of a
// 256-byte window, so reading a byte is a readlane.  Offsets are 32-bit and
// relative to the tile's first byte (the window base may lie below it).  The
// next lower window is loaded when the reader crosses into it and waited for
// there: a load kept in flight across the symbol loop would make every
// loop-carried register copy wait for it, and in-order completion would also
// make it wait for every coefficient store before it.
struct ByteWin {
  const uint8_t* tile;  // tile byte 0
  int32_t lo, hi;       // dword-aligned readable range [lo, hi), tile-relative
  int32_t wbase;        // current window base, tile-relative (tile + wbase is 4-aligned)
  int32_t lim;          // max(wbase, 0): bytes at positions >= lim may be read from the window
  uint32_t cur;

  __device__ __forceinline__ void load(int lane) {
    const int32_t a = wbase + 4 * lane;
    using GWord = __attribute__((address_space(1))) const uint32_t;  // packets live in device global memory
    // The address is clamped into the readable range so the load needs no
    // branch; lanes outside it then read as zero.
    const int32_t ac = a < lo ? lo : (a >= hi ? hi - 4 : a);
    const uint32_t v = *(GWord*)(tile + ac);
    cur = (a >= lo && a < hi) ? v : 0u;
    lim = wbase > 0 ? wbase : 0;
  }
  __device__ void init(const uint8_t* t, uint32_t len, int lane) {
    tile = t;
    const int32_t mis = static_cast<int32_t>(reinterpret_cast<uintptr_t>(t) & 255u);
    lo = -static_cast<int32_t>(reinterpret_cast<uintptr_t>(t) & 3u);
    hi = (static_cast<int32_t>(len) - lo + 3) / 4 * 4 + lo;
    wbase = ((static_cast<int32_t>(len ? len - 1 : 0) + mis) & ~255) - mis;
    load(lane);
  }
  // Byte p of the tile; p only ever decreases between calls.
  __device__ __forceinline__ uint32_t byte(int32_t p, int lane) {
    if (p < wbase) {
      wbase -= 256;
      load(lane);
    }
    const uint32_t off = static_cast<uint32_t>(p - wbase);
    return (rd_lane(cur, static_cast<int>(off >> 2)) >> ((off & 3u) * 8u)) & 0xFFu;
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
// The tile's MV rows live in LDS; typed as LDS pointers their accesses stay
// LDS instructions (generic pointers would make them flat accesses, and a flat
// load waits for every outstanding store).
using LdsU32 = __attribute__((address_space(3))) uint32_t;

__device__ __forceinline__ uint32_t predict_mv(const LdsU32* prow, const LdsU32* crow, int lx, int tw, bool top_row) {
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

// ----------------------------------------------------------------------------
// Entropy decoding of one tile.  All state is wave-uniform; the lanes take
// part only in the symbol search (one lane per table entry), the byte window
// and the coefficient staging buffer.
//
// Symbol search.  The frame's tables sit in 15 VGPRs (FrameHdr::vt): lane l
// of a vector holds one packed entry start << 16 | freq.  For a slot the
// lanes of the context's group whose start (the upper half, compared in
// place with an SDWA word select) is at most the slot are counted; the count
// minus one is the symbol (the last symbol whose range starts at or below
// the slot, rans.go:120-148 slot2sym), and a count of zero means an unused
// context (every entry is kNoSym, whose start 0xFFFF exceeds every slot).
// One compare over the vector, one ballot, one popcount and one readlane
// (start and frequency together) decode a symbol; no memory access is on the
// dependency chain from one symbol to the next.
// ----------------------------------------------------------------------------
// Common tail of the single-symbol decoders: xn is the updated state.
// Renormalise from the lane window while pos > lim (lim = max(window base, 0)),
// then rotate the states (rans.go:286-303).  st = 0 done, 1 unused context,
// 2 renormalisation to be completed by advance() from xn.
#define EBIV_DECODE_TAIL                               \
  "s_cmp_lt_u32 %[xn], 0x800000\n"                    \
  "s_cbranch_scc1 L%=_renorm\n"                       \
  "L%=_rot:\n"                                        \
  "s_mov_b32 %[x0], %[x1]\n"                          \
  "s_mov_b32 %[x1], %[x2]\n"                          \
  "s_mov_b32 %[x2], %[x3]\n"                          \
  "s_mov_b32 %[x3], %[xn]\n"                          \
  "s_add_u32 %[nsym], %[nsym], 1\n"                   \
  "s_mov_b32 %[st], 0\n"                              \
  "s_branch L%=_end\n"                                \
  "L%=_renorm:\n"                                     \
  "s_cmp_le_i32 %[pos], %[lim]\n"                     \
  "s_cbranch_scc1 L%=_slow\n"                         \
  "s_sub_u32 %[pos], %[pos], 1\n"                     \
  "s_sub_u32 %[t1], %[pos], %[wbase]\n"               \
  "s_lshr_b32 %[t2], %[t1], 2\n"                      \
  "v_readlane_b32 %[t2], %[vwin], %[t2]\n"            \
  "s_and_b32 %[t1], %[t1], 3\n"                       \
  "s_lshl_b32 %[t1], %[t1], 3\n"                      \
  "s_lshr_b32 %[t2], %[t2], %[t1]\n"                  \
  "s_and_b32 %[t2], %[t2], 0xff\n"                    \
  "s_lshl_b32 %[xn], %[xn], 8\n"                      \
  "s_or_b32 %[xn], %[xn], %[t2]\n"                    \
  "s_cmp_lt_u32 %[xn], 0x800000\n"                    \
  "s_cbranch_scc1 L%=_renorm\n"                       \
  "s_branch L%=_rot\n"                                \
  "L%=_slow:\n"                                       \
  "s_mov_b32 %[st], 2\n"                              \
  "s_branch L%=_end\n"                                \
  "L%=_unused:\n"                                     \
  "s_mov_b32 %[st], 1\n"                              \
  "L%=_end:\n"

struct Rans {
  uint32_t x0, x1, x2, x3;  // x0 is state[counter & 3] (rans.go:279-282)
  uint32_t pos;             // tile bytes not yet consumed
  uint32_t nsym;
  int32_t err;
  ByteWin win;

  // Renormalisation and state rotation after a symbol (rans.go:286-303).
  // Most symbols need no byte, so the byte loop is the rare branch.
  __device__ __forceinline__ bool advance(uint32_t x, int lane) {
    if (__builtin_expect(x < kRansLower, 0)) {
      do {
        if (pos == 0) {
          err = kErrCorrupt | (kTileUnderrun << 8);
          return false;
        }
        x = (x << 8) | win.byte(static_cast<int32_t>(--pos), lane);
      } while (x < kRansLower);
    }
    x0 = x1; x1 = x2; x2 = x3; x3 = x;
    nsym++;
    return true;
  }
  // ransDecoder.decode (rans.go:270-304) for the context whose entries are
  // the lanes [goff, goff + width) of V.  Returns the symbol, or -1 on error.
  // One hand-scheduled asm statement: table compare, ballot, popcount, lane
  // read, state update and (when the window holds the bytes) the
  // renormalisation; a renormalisation that reaches the window's base or the
  // stream's start is completed by advance().
  template <uint32_t kGoff, uint32_t kWidth>
  __device__ __forceinline__ int decode(uint32_t V, int lane) {
    static_assert(kGoff + kWidth <= 64 && kWidth <= 32, "lane group");
    int32_t st, sym;
    uint32_t xn, slot, q, t1, t2, p;
    uint64_t b;
    asm volatile(
        "s_and_b32 %[slot], %[x0], 0xfff\n"
        "v_cmp_ge_u32_sdwa %[b], %[slot], %[V] src0_sel:DWORD src1_sel:WORD_1\n"
        "s_lshr_b32 %[q], %[x0], 12\n"
        "s_bfe_u64 %[b], %[b], %[bfe]\n"
        "s_bcnt1_i32_b64 %[sym], %[b]\n"
        "s_cmp_eq_u32 %[sym], 0\n"
        "s_cbranch_scc1 L%=_unused\n"
        "s_add_u32 %[t1], %[sym], %[gm1]\n"
        "v_readlane_b32 %[p], %[V], %[t1]\n"
        "s_add_u32 %[sym], %[sym], -1\n"
        "s_and_b32 %[t2], %[p], 0xffff\n"
        "s_lshr_b32 %[p], %[p], 16\n"
        "s_mul_i32 %[xn], %[t2], %[q]\n"
        "s_sub_u32 %[t2], %[slot], %[p]\n"
        "s_add_u32 %[xn], %[xn], %[t2]\n"
        EBIV_DECODE_TAIL
        : [x0] "+&s"(x0), [x1] "+&s"(x1), [x2] "+&s"(x2), [x3] "+&s"(x3), [pos] "+&s"(pos), [nsym] "+&s"(nsym),
          [st] "=&s"(st), [sym] "=&s"(sym), [xn] "=&s"(xn), [slot] "=&s"(slot), [q] "=&s"(q), [t1] "=&s"(t1),
          [t2] "=&s"(t2), [p] "=&s"(p), [b] "=&s"(b)
        : [V] "v"(V), [bfe] "i"(kGoff | (kWidth << 16)), [gm1] "i"(kGoff - 1), [lim] "s"(win.lim),
          [wbase] "s"(win.wbase), [vwin] "v"(win.cur)
        : "scc");
    if (__builtin_expect(st != 0, 0)) {
      if (st == 1) {
        err = kErrCorrupt | (kTileUnused << 8);
        return -1;
      }
      if (!advance(xn, lane)) return -1;
    }
    return sym;
  }
  // decode() with a lane group known only at run time (the fallback token
  // loop).  Returns the symbol, or -1 on error.
  __device__ __forceinline__ int decode_rt(uint32_t V, uint32_t goff, uint32_t gmask, int lane) {
    const uint32_t x = x0;
    const uint32_t slot = x & (kRansM - 1);
    const uint64_t b = __ballot((V >> 16) <= slot);
    const uint32_t cnt = static_cast<uint32_t>(__builtin_popcount(static_cast<uint32_t>(b >> goff) & gmask));
    if (__builtin_expect(cnt == 0, 0)) {
      err = kErrCorrupt | (kTileUnused << 8);
      return -1;
    }
    const uint32_t p = rd_lane(V, static_cast<int>(goff + cnt - 1));
    if (!advance((p & 0xFFFFu) * (x >> kRansScaleBits) + slot - (p >> 16), lane)) return -1;
    return static_cast<int>(cnt - 1);
  }
  // The bypass context, fixed at {2048, 2048} (rans.go:312-322).
  __device__ __forceinline__ int bit(int lane) {
    int32_t st, sym;
    uint32_t xn, t1, t2;
    asm volatile(
        "s_lshr_b32 %[xn], %[x0], 12\n"
        "s_lshl_b32 %[xn], %[xn], 11\n"
        "s_and_b32 %[t1], %[x0], 0x7ff\n"
        "s_bfe_u32 %[sym], %[x0], 0x1000b\n"
        "s_add_u32 %[xn], %[xn], %[t1]\n"
        EBIV_DECODE_TAIL
        : [x0] "+&s"(x0), [x1] "+&s"(x1), [x2] "+&s"(x2), [x3] "+&s"(x3), [pos] "+&s"(pos), [nsym] "+&s"(nsym),
          [st] "=&s"(st), [sym] "=&s"(sym), [xn] "=&s"(xn), [t1] "=&s"(t1), [t2] "=&s"(t2)
        : [lim] "s"(win.lim), [wbase] "s"(win.wbase), [vwin] "v"(win.cur)
        : "scc");
    if (__builtin_expect(st != 0, 0)) {
      if (!advance(xn, lane)) return -1;
    }
    return sym;
  }
  // getClassValue (mv.go:95-108): a class symbol, then class-1 raw bits under
  // an implicit leading one.  Returns -1 on error.
  template <uint32_t kGoff>
  __device__ __forceinline__ int class_value(uint32_t V, int lane) {
    const int v = decode<kGoff, 16>(V, lane);
    if (v <= 1) return v;
    int val = 1;
    for (int i = 1; i < v; i++) {
      const int b = bit(lane);
      if (b < 0) return -1;
      val = (val << 1) | b;
    }
    return val;
  }
};

// The table vectors of one tile's wavefront.
struct Tables {
  uint32_t misc0, misc1, misc2;  // non-token contexts
  // token contexts t<set><band pair>: set = luma tx 0-2 or 3 for chroma.
  // Named registers rather than an array: an array indexed by the transform
  // size would be placed in scratch memory.
  uint32_t t00, t01, t02, t10, t11, t12, t20, t21, t22, t30, t31, t32;
};
// Lane groups of the non-token contexts (vector, first lane).
constexpr uint32_t kGSign = 0, kGMBMode = 16, kGTx = 32, kGCBP = 48;       // misc0
constexpr uint32_t kGLumaMode = 0, kGChromaMode = 16, kGRef = 32, kGPart = 48;  // misc1
constexpr uint32_t kGEsc = 0, kGMVX = 16, kGMVY = 32;                       // misc2
static_assert(kSignVec == 0 && kGSign == kSignLane, "sign table position");

// Coefficient list output: lane k stages entry base+k until 64 are queued.
struct CoefOut {
  uint32_t* dst;
  uint32_t base, cn;  // entries stored, entries staged
  uint32_t buf;
  __device__ __forceinline__ void emit(uint32_t e, int lane) {
    if (static_cast<uint32_t>(lane) == cn) buf = e;
    if (++cn == 64) {
      dst[base + lane] = buf;
      base += 64;
      cn = 0;
    }
  }
};

// decodeBlock (residual.go:79-126) for one block: tokens in zigzag order under
// (plane, size, band, previous class) contexts, escapes, explicit signs, and
// the first sign hidden in the level-sum parity when the span allows.  Levels
// go to the coefficient list; the first coefficient goes last, once its sign
// is known.  va/vb/vc are the token vectors of the block's (plane, size) for
// bands 0-1, 2-3 and 4; the context (band, prev) is the lane group
// 8 * ((band & 1) * 3 + prev) of its band's vector.  Returns false on a decode
// error.
__device__ __forceinline__ bool parse_block(Rans& R, CoefOut& out, const Tables& T, uint32_t va, uint32_t vb,
                                            uint32_t vc, uint32_t zzv, int total, int blk, int lane) {
  uint32_t vcur = va, boff = 0, prevoff = 0;
  int si = 0, band = 0, bnext = 1, first = -1, last = -1, first_pos = 0;
  int32_t sum_abs = 0, first_mag = 0;
  for (;;) {
    const int t = R.decode_rt(vcur, boff + prevoff, 0xFFu, lane);
    if (t < 0) return false;
    if (t == kTEOB) break;
    if (t == kTZero) {
      prevoff = 0;
    } else {
      int32_t mag = t;
      prevoff = t == kTOne ? 8u : 16u;  // classForMag: 1 -> 1, else 2 (model.go:181-191)
      if (t == kTEscape) {
        const int e = R.class_value<kGEsc>(T.misc2, lane);
        if (e < 0) return false;
        mag = 5 + e;
      }
      sum_abs += mag;
      if (first < 0) {
        first = si;
        first_mag = mag;
        first_pos = si;
      } else {
        const int s = R.decode<kGSign, 16>(T.misc0, lane);
        if (s < 0) return false;
        out.emit(coef_entry(si, blk, s ? -mag : mag), lane);
      }
      last = si;
    }
    if (++si == total) break;
    if (si == bnext) {  // bandOf thresholds 1, 3, 9, 21 (model.go:165-180)
      band++;
      boff ^= 24u;  // (band & 1) * 3 rows of 8 lanes
      if (band == 2) vcur = vb;
      if (band == 4) vcur = vc;
      bnext = band < 4 ? static_cast<int>((0x15090301u >> (8 * band)) & 0xFFu) : (1 << 20);
    }
  }
  if (first >= 0) {
    bool neg = (sum_abs & 1) != 0;
    if (last - first < kSDHMinSpan) {
      const int s = R.decode<kGSign, 16>(T.misc0, lane);
      if (s < 0) return false;
      neg = s != 0;
    }
    out.emit(coef_entry(first_pos, blk, neg ? -first_mag : first_mag), lane);
  }
  return true;
}

// Grammar of one macroblock as decoded so far: what the record stores.
struct MBState {
  uint32_t kind, part, cbp, tx, refsel, cmode, wmask, lmode0, lmode1;
  uint32_t mv0, mv1, mv2, mv3, mvcur;
};

// A macroblock after its mode, as one pass over its blocks so that the block
// decoder is inlined once.
//   intra (decodeLumaMB / decodeChromaMB, decode_intra.go:137-167): transform
//     size, then per luma block its mode and residual, then the chroma mode
//     and both chroma residuals;
//   inter (decodeInterMB / reconInterMB, decode_inter.go:18-118): reference,
//     partition, one MV per partition (the first against the predictor, the
//     others against the first), the coded-block pattern, then the coded
//     residuals (transform size first when luma is coded).
// The write mask follows what Go writes before an error stops the tile: a
// block once its residual decoded; uncoded inter luma once the pattern is
// known; uncoded chroma planes once luma is complete (even when an earlier
// coded chroma plane then fails); nothing when the failure comes before the
// pattern.
__device__ __forceinline__ bool parse_mb(Rans& R, CoefOut& out, MBState& m, const Tables& T, uint32_t zz4, uint32_t zz8, uint32_t zz16, bool intra, uint32_t pred_last,
                                         int lane) {
  bool luma_coded = true, cb_coded = true, cr_coded = true;
  if (!intra) {
    const int refsel = R.decode<kGRef, 16>(T.misc1, lane);
    if (refsel < 0) return false;
    m.refsel = static_cast<uint32_t>(refsel);
    const uint32_t pred = refsel == kRefGolden ? 0u : pred_last;
    const int part = R.decode<kGPart, 16>(T.misc1, lane);
    if (part < 0) return false;
    m.part = static_cast<uint32_t>(part);
    const int npart = static_cast<int>((kPartCountNibbles >> (4 * part)) & 15u);
    for (int pi = 0; pi < npart; pi++) {
      const uint32_t base = pi > 0 ? m.mv0 : pred;
      uint32_t mv = 0;
      for (int comp = 0; comp < 2; comp++) {
        const int v = comp ? R.class_value<kGMVY>(T.misc2, lane) : R.class_value<kGMVX>(T.misc2, lane);
        if (v < 0) return false;
        mv |= static_cast<uint32_t>(static_cast<uint16_t>(clamp_mv(mv_comp(base, comp) + zigzag_dec(static_cast<uint32_t>(v)))))
              << (16 * comp);
      }
      if (pi == 0) m.mv0 = mv;
      else if (pi == 1) m.mv1 = mv;
      else if (pi == 2) m.mv2 = mv;
      else m.mv3 = mv;
    }
    m.mvcur = refsel == kRefGolden ? 0u : m.mv0;
    const int cbp = R.decode<kGCBP, 16>(T.misc0, lane);
    if (cbp < 0) return false;
    m.cbp = static_cast<uint32_t>(cbp);
    luma_coded = (cbp & kCBPLuma) != 0;
    cb_coded = (cbp & kCBPCb) != 0;
    cr_coded = (cbp & kCBPCr) != 0;
    if (!luma_coded) m.wmask |= 0xFFFFu;
  }
  int nb = 0, total = 0, tx = 0;
  uint32_t zzv = zz4;
  if (luma_coded) {
    tx = R.decode<kGTx, 16>(T.misc0, lane);
    if (tx < 0) return false;
    m.tx = static_cast<uint32_t>(tx);
    const int n = 4 << tx;
    nb = (16 / n) * (16 / n);
    total = n * n;
    zzv = tx == 0 ? zz4 : (tx == 1 ? zz8 : zz16);
  }
  // The luma blocks' token vectors (set = transform index).
  // The empty asm hands over the registers as opaque values: a selection
  // between fields of one struct is otherwise turned into an indexed load,
  // which puts the struct in scratch memory.
  uint32_t a0 = T.t00, a1 = T.t10, a2 = T.t20, b0 = T.t01, b1 = T.t11, b2 = T.t21, c0 = T.t02, c1 = T.t12, c2 = T.t22;
  uint32_t ca = T.t30, cb = T.t31, cc = T.t32;
  asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(b0), "+v"(b1), "+v"(b2));
  asm volatile("" : "+v"(c0), "+v"(c1), "+v"(c2), "+v"(ca), "+v"(cb), "+v"(cc));
  const uint32_t la = tx == 0 ? a0 : (tx == 1 ? a1 : a2);
  const uint32_t lb = tx == 0 ? b0 : (tx == 1 ? b1 : b2);
  const uint32_t lc = tx == 0 ? c0 : (tx == 1 ? c1 : c2);
  for (int k = 0; k < nb + 2; k++) {
    const bool chroma = k >= nb;
    const int blk = chroma ? 16 + (k - nb) : k;
    if (k == nb) {  // luma complete
      if (intra) {
        const int cmode = R.decode<kGChromaMode, 16>(T.misc1, lane);
        if (cmode < 0) return false;
        m.cmode = static_cast<uint32_t>(cmode);
      } else {
        if (!cb_coded) m.wmask |= 1u << 16;
        if (!cr_coded) m.wmask |= 1u << 17;
      }
    }
    if (chroma && !(blk == 16 ? cb_coded : cr_coded)) continue;
    if (intra && !chroma) {
      const int mode = R.decode<kGLumaMode, 16>(T.misc1, lane);
      if (mode < 0) return false;
      if (k < 8) m.lmode0 |= static_cast<uint32_t>(mode) << (4 * k);
      else m.lmode1 |= static_cast<uint32_t>(mode) << (4 * (k - 8));
    }
    const uint32_t va = chroma ? ca : la, vb = chroma ? cb : lb, vc = chroma ? cc : lc;
    const uint32_t zb = chroma ? zz8 : zzv;
    const int tb = chroma ? 64 : total;
    if (!parse_block(R, out, T, va, vb, vc, zb, tb, blk, lane)) return false;
    m.wmask |= 1u << blk;
  }
  return true;
}

// ============================================================================
// Entropy kernel, hand-scheduled macroblock loop
// ============================================================================
// decodeTile (decode_intra.go:121-135) as one asm statement: macroblock
// modes, MV prediction (mv.go:21-52), the inter header (decode_inter.go:18-62:
// reference, partition, MV class values with raw bits, coded-block pattern),
// transform size and intra modes (decode_intra.go:137-167), every token of
// every block with escapes, explicit and hidden signs (residual.go:79-126),
// coefficient entries, macroblock records and the MV rows.  All decoder state
// is uniform and lives in SGPRs; vector registers hold the symbol tables, the
// byte window, the scan tables, the MV rows (lane = tile column) and two
// staging registers (64 coefficient entries; 5 records of 12 dwords).
//
// It returns at the end of the tile (kMbDone) or at a decode error
// (kMbUnused, kMbUnderrun) with the failing macroblock's fields decoded so
// far, which the caller turns into its record exactly as the generic path
// does.  Precondition: the sign context has a table (signs from two SGPRs).
enum MbEvent : uint32_t { kMbDone = 0, kMbUnused = 3, kMbUnderrun = 4 };

struct MbLoop {
  // rANS, window, outputs (in/out)
  uint32_t x0, x1, x2, x3, pos, nsym, wbase, lim, cn, cbase;
  uint32_t mbx, mby, done, nintra, leftmv;
  uint32_t vwin, vbuf;    // per lane
  uint32_t prowa, crowa;  // LDS MV rows: the row above, the current row
  // the failing macroblock (outputs)
  uint32_t r0, r1, wmask, cstart, mv0, mv1, mv2, mv3, mvcur, lm0, lm1;
};

// Cost model of one wavefront alone on a SIMD (gfx950, measured): an SALU
// instruction issues every 4 cycles; after a VALU instruction that writes an
// SGPR (compare to a mask, readlane) the next SALU instruction waits about
// 20 cycles, which VALU instructions can fill but SALU ones cannot; a
// conditional branch costs about 10 cycles not taken and 25 taken.  The loop
// is laid out so that the common paths fall through, and the VALU work of a
// step (symbol search, the speculative table read of symbol 0, the MV
// predictor) is issued together ahead of one wait.

// Register roles that share one operand (never live together).
#define MB_MAG "%[val]"
#define MB_NEG "%[k2]"
#define MB_MVT "%[sumabs]"
#define MB_PI "%[si]"
#define MB_NPART "%[stop]"
#define MB_PRED "%[firstpm]"

// Renormalisation after a symbol (xn = updated state), in line: branch out
// when a byte is needed.  Out of line (MB_RN_OOL): the byte loop with
// underrun check and window slide, then back to the rotation.
#define MB_RN(tag)                               \
  "s_cmp_lt_u32 %[xn], 0x800000\n"               \
  "s_cbranch_scc1 L%=_" tag "_rn\n"              \
  "L%=_" tag "_rot:\n"                           \
  "s_mov_b32 %[x0], %[x1]\n"                     \
  "s_mov_b32 %[x1], %[x2]\n"                     \
  "s_mov_b32 %[x2], %[x3]\n"                     \
  "s_mov_b32 %[x3], %[xn]\n"                     \
  "s_add_u32 %[nsym], %[nsym], 1\n"
#define MB_RN_OOL(tag, underrun)                 \
  "L%=_" tag "_rn:\n"                            \
  "s_cmp_eq_u32 %[pos], 0\n"                     \
  "s_cbranch_scc1 L%=_" underrun "\n"            \
  "s_sub_u32 %[pos], %[pos], 1\n"                \
  "s_cmp_lt_i32 %[pos], %[wbase]\n"              \
  "s_cbranch_scc1 L%=_" tag "_sl\n"              \
  "L%=_" tag "_by:\n"                            \
  "s_sub_u32 %[t1], %[pos], %[wbase]\n"          \
  "s_lshr_b32 %[t2], %[t1], 2\n"                 \
  "v_readlane_b32 %[t2], %[vwin], %[t2]\n"       \
  "s_and_b32 %[t1], %[t1], 3\n"                  \
  "s_lshl_b32 %[t1], %[t1], 3\n"                 \
  "s_lshr_b32 %[t2], %[t2], %[t1]\n"             \
  "s_and_b32 %[t2], %[t2], 0xff\n"               \
  "s_lshl_b32 %[xn], %[xn], 8\n"                 \
  "s_or_b32 %[xn], %[xn], %[t2]\n"               \
  "s_cmp_lt_u32 %[xn], 0x800000\n"               \
  "s_cbranch_scc1 L%=_" tag "_rn\n"              \
  "s_branch L%=_" tag "_rot\n"                   \
  "L%=_" tag "_sl:\n"                            \
  "s_sub_u32 %[wbase], %[wbase], 256\n"          \
  "s_max_i32 %[lim], %[wbase], 0\n"              \
  "s_sub_u32 %[t1], %[wbase], %[wlo]\n"          \
  "v_add_u32 %[vaddr], %[t1], %[vlane4]\n"       \
  "v_max_i32 %[vt], %[vaddr], 0\n"               \
  "v_min_i32 %[vt], %[hi2], %[vt]\n"             \
  "global_load_dword %[vwin], %[vt], %[tb]\n"    \
  "v_cmp_ge_u32_e32 vcc, %[hi2], %[vaddr]\n"     \
  "s_waitcnt vmcnt(0)\n"                         \
  "v_cndmask_b32_e32 %[vwin], 0, %[vwin], vcc\n" \
  "s_branch L%=_" tag "_by\n"

// Symbol of the context at lanes [goff, goff + 16) of vec into dst.  The
// compare takes the entries' start (upper half) against the slot.
#define MB_DEC(vec, goff, gm1, bfe, dst, tag, unused) \
  "s_and_b32 %[slot], %[x0], 0xfff\n"            \
  "v_cmp_ge_u32_sdwa %[b], %[slot], " vec " src0_sel:DWORD src1_sel:WORD_1\n" \
  "s_lshr_b32 %[xn], %[x0], 12\n"                \
  "s_bfe_u64 %[b], %[b], " bfe "\n"              \
  "s_bcnt1_i32_b64 " dst ", %[b]\n"              \
  "s_cmp_eq_u32 " dst ", 0\n"                    \
  "s_cbranch_scc1 L%=_" unused "\n"              \
  "s_add_u32 %[t1], " dst ", " gm1 "\n"          \
  "v_readlane_b32 %[p], " vec ", %[t1]\n"        \
  "s_add_u32 " dst ", " dst ", -1\n"             \
  "s_and_b32 %[t2], %[p], 0xffff\n"              \
  "s_lshr_b32 %[p], %[p], 16\n"                  \
  "s_mul_i32 %[xn], %[t2], %[xn]\n"              \
  "s_sub_u32 %[t2], %[slot], %[p]\n"             \
  "s_add_u32 %[xn], %[xn], %[t2]\n"              \
  MB_RN(tag)

// A raw bit (the bypass context {2048, 2048}) into dst.
#define MB_BIT(dst, tag)                         \
  "s_lshr_b32 %[xn], %[x0], 12\n"                \
  "s_lshl_b32 %[xn], %[xn], 11\n"                \
  "s_and_b32 %[t1], %[x0], 0x7ff\n"              \
  "s_bfe_u32 " dst ", %[x0], 0x1000b\n"          \
  "s_add_u32 %[xn], %[xn], %[t1]\n"              \
  MB_RN(tag)

// getClassValue (mv.go:95-108) of the 16-symbol context at goff of vec into
// %[sym]: class, then class-1 raw bits under a leading one.
#define MB_CLASS(vec, goff, gm1, bfe, tag)       \
  MB_DEC(vec, goff, gm1, bfe, "%[sym]", tag "d", "unused") \
  "s_cmp_lt_u32 %[sym], 2\n"                     \
  "s_cbranch_scc1 L%=_" tag "_cv\n"              \
  "s_mov_b32 %[val], 1\n"                        \
  "s_add_u32 %[k2], %[sym], -1\n"                \
  "L%=_" tag "_bits:\n"                          \
  MB_BIT("%[t3]", tag "b")                       \
  "s_lshl_b32 %[val], %[val], 1\n"               \
  "s_or_b32 %[val], %[val], %[t3]\n"             \
  "s_add_u32 %[k2], %[k2], -1\n"                 \
  "s_cmp_lg_u32 %[k2], 0\n"                      \
  "s_cbranch_scc1 L%=_" tag "_bits\n"            \
  "s_mov_b32 %[sym], %[val]\n"                   \
  "L%=_" tag "_cv:\n"

// Explicit sign from the sign context's two frequencies: mag <- -mag on 1.
#define MB_SIGN(tag)                             \
  "s_sub_u32 " MB_NEG ", 0, " MB_MAG "\n"        \
  "s_and_b32 %[slot], %[x0], 0xfff\n"            \
  "s_lshr_b32 %[xn], %[x0], 12\n"                \
  "s_cmp_ge_u32 %[slot], %[f0]\n"                \
  "s_cselect_b32 %[p], %[f0], 0\n"               \
  "s_cselect_b32 %[t2], %[f1], %[f0]\n"          \
  "s_cselect_b32 " MB_MAG ", " MB_NEG ", " MB_MAG "\n" \
  "s_mul_i32 %[xn], %[t2], %[xn]\n"              \
  "s_sub_u32 %[t2], %[slot], %[p]\n"             \
  "s_add_u32 %[xn], %[xn], %[t2]\n"              \
  MB_RN(tag)

// Coefficient entry (scan index POS, level mag) into the staging register;
// 64 staged entries go out in one store (out of line: MB_EMIT_OOL).
#define MB_EMIT(tag) MB_EMIT_AT(tag, "%[t2]")
#define MB_EMIT_AT(tag, POS)                     \
  "s_lshl_b32 %[t1], " MB_MAG ", 13\n"           \
  "s_or_b32 %[t1], %[t1], " POS "\n"             \
  "s_or_b32 %[t1], %[t1], %[blk8]\n"             \
  "s_mov_b32 m0, %[cn]\n"                        \
  "v_writelane_b32 %[vbuf], %[t1], m0\n"         \
  "s_add_u32 %[cn], %[cn], 1\n"                  \
  "s_cmp_eq_u32 %[cn], 64\n"                     \
  "s_cbranch_scc1 L%=_" tag "_cf\n"              \
  "L%=_" tag "_cfd:\n"
#define MB_EMIT_OOL(tag)                         \
  "L%=_" tag "_cf:\n"                            \
  "s_lshl_b32 %[t1], %[cbase], 2\n"              \
  "v_add_u32 %[vaddr], %[t1], %[vlane4]\n"       \
  "global_store_dword %[vaddr], %[vbuf], %[dst]\n" \
  "s_add_u32 %[cbase], %[cbase], 64\n"           \
  "s_mov_b32 %[cn], 0\n"                         \
  "s_branch L%=_" tag "_cfd\n"

// ---- token loop specialised by rANS phase ----------------------------------
// Four copies of the hot token loop, one per interleave phase P: the next
// symbol uses state register X[P] (x0..x3), which the decode updates in
// place, so a symbol costs no register rotation.  A zero token (the common
// case) falls through into the next phase's copy; other tokens leave the
// straight line.  Leaving the loop (escape, end of block) rotates the
// registers back to phase 0, where the rest of the loop keeps x0 as the next
// symbol's state.
//   off  = lane of the token context's symbol 0 (band group + previous
//          class: 0, 8, 16); offz the same with previous class 0;
//   bf   = the field of the context in the ballot (width 8 at off), and bfz.
#define MBP_RN(tag, R)                           \
  "s_cmp_lt_u32 " R ", 0x800000\n"               \
  "s_cbranch_scc1 L%=_" tag "_rn\n"              \
  "L%=_" tag "_rot:\n"                           \
  "s_add_u32 %[nsym], %[nsym], 1\n"
#define MBP_RN_OOL(tag, underrun, R)             \
  "L%=_" tag "_rn:\n"                            \
  "s_cmp_eq_u32 %[pos], 0\n"                     \
  "s_cbranch_scc1 L%=_" underrun "\n"            \
  "s_sub_u32 %[pos], %[pos], 1\n"                \
  "s_cmp_lt_i32 %[pos], %[wbase]\n"              \
  "s_cbranch_scc1 L%=_" tag "_sl\n"              \
  "L%=_" tag "_by:\n"                            \
  "s_sub_u32 %[t1], %[pos], %[wbase]\n"          \
  "s_lshr_b32 %[t2], %[t1], 2\n"                 \
  "v_readlane_b32 %[t2], %[vwin], %[t2]\n"       \
  "s_and_b32 %[t1], %[t1], 3\n"                  \
  "s_lshl_b32 %[t1], %[t1], 3\n"                 \
  "s_lshr_b32 %[t2], %[t2], %[t1]\n"             \
  "s_and_b32 %[t2], %[t2], 0xff\n"               \
  "s_lshl_b32 " R ", " R ", 8\n"                 \
  "s_or_b32 " R ", " R ", %[t2]\n"               \
  "s_cmp_lt_u32 " R ", 0x800000\n"               \
  "s_cbranch_scc1 L%=_" tag "_rn\n"              \
  "s_branch L%=_" tag "_rot\n"                   \
  "L%=_" tag "_sl:\n"                            \
  "s_sub_u32 %[wbase], %[wbase], 256\n"          \
  "s_max_i32 %[lim], %[wbase], 0\n"              \
  "s_sub_u32 %[t1], %[wbase], %[wlo]\n"          \
  "v_add_u32 %[vaddr], %[t1], %[vlane4]\n"       \
  "v_max_i32 %[vt], %[vaddr], 0\n"               \
  "v_min_i32 %[vt], %[hi2], %[vt]\n"             \
  "global_load_dword %[vwin], %[vt], %[tb]\n"    \
  "v_cmp_ge_u32_e32 vcc, %[hi2], %[vaddr]\n"     \
  "s_waitcnt vmcnt(0)\n"                         \
  "v_cndmask_b32_e32 %[vwin], 0, %[vwin], vcc\n" \
  "s_branch L%=_" tag "_by\n"
// As MBP_RN, without counting the symbol: the token symbols of a block are
// counted once at its end (one per scan position passed, plus the end of
// block token).
#define MBP_RN_NC(tag, R)                        \
  "s_cmp_lt_u32 " R ", 0x800000\n"               \
  "s_cbranch_scc1 L%=_" tag "_rn\n"              \
  "L%=_" tag "_rot:\n"
// Explicit sign (as MB_SIGN) with state register R updated in place.
#define MBP_SIGN(tag, R)                         \
  "s_sub_u32 " MB_NEG ", 0, " MB_MAG "\n"        \
  "s_and_b32 %[slot], " R ", 0xfff\n"            \
  "s_lshr_b32 " R ", " R ", 12\n"                \
  "s_cmp_ge_u32 %[slot], %[f0]\n"                \
  "s_cselect_b32 %[p], %[f0], 0\n"               \
  "s_cselect_b32 %[t2], %[f1], %[f0]\n"          \
  "s_cselect_b32 " MB_MAG ", " MB_NEG ", " MB_MAG "\n" \
  "s_mul_i32 " R ", %[t2], " R "\n"              \
  "s_sub_u32 %[t2], %[slot], %[p]\n"             \
  "s_add_u32 " R ", " R ", %[t2]\n"              \
  MBP_RN(tag, R)
// The token at phase P (state RA) in state F: "a" before the block's first
// level (whose sign resolves at the end of the block), "b" after it.  The
// straight line is the zero token: the compare and the table entry of
// symbol 0 (start 0, so the entry is its frequency) are read together; it
// falls through into phase NP1.
#define MBP_TOK(P, NP1, RA, F)                   \
  "L%=_tok" P #F ":\n"                           \
  "s_and_b32 %[slot], " RA ", 0xfff\n"           \
  "v_cmp_ge_u32_sdwa %[b], %[slot], %[vcur] src0_sel:DWORD src1_sel:WORD_1\n" \
  "v_readlane_b32 %[p], %[vcur], %[off]\n"       \
  "s_lshr_b32 " RA ", " RA ", 12\n"              \
  "s_bfe_u64 %[b], %[b], %[bf]\n"                \
  "s_cmp_eq_u64 %[b], 1\n"                       \
  "s_cbranch_scc0 L%=_tnz" P #F "\n"             \
  "s_mul_i32 " RA ", %[p], " RA "\n"             \
  "s_add_u32 " RA ", " RA ", %[slot]\n"          \
  MBP_RN_NC("tz" P #F, RA)                       \
  "s_mov_b32 %[off], %[offz]\n"                  \
  "s_mov_b32 %[bf], %[bfz]\n"                    \
  "s_add_u32 %[si], %[si], 1\n"                  \
  "s_cmp_eq_u32 %[si], %[stop]\n"                \
  "s_cbranch_scc1 L%=_tstop" NP1 #F "\n"
// Any other token at phase P (out of line).  RB is the next phase's state
// (the sign's); NP1, NP2: the phases after one and two symbols.
#define MBP_NZ(P, NP1, NP2, RA, RB, F)           \
  "L%=_tnz" P #F ":\n"                           \
  "s_bcnt1_i32_b64 %[cnt], %[b]\n"               \
  "s_cmp_eq_u32 %[cnt], 0\n"                     \
  "s_cbranch_scc1 L%=_unused\n"                  \
  "s_add_u32 %[t1], %[off], %[cnt]\n"            \
  "s_add_u32 %[t1], %[t1], -1\n"                 \
  "v_readlane_b32 %[p], %[vcur], %[t1]\n"        \
  "s_add_u32 " MB_MAG ", %[cnt], -1\n"           \
  "s_and_b32 %[t2], %[p], 0xffff\n"              \
  "s_lshr_b32 %[p], %[p], 16\n"                  \
  "s_mul_i32 " RA ", %[t2], " RA "\n"            \
  "s_sub_u32 %[t2], %[slot], %[p]\n"             \
  "s_add_u32 " RA ", " RA ", %[t2]\n"            \
  MBP_RN_NC("tn" P #F, RA)                       \
  "s_cmp_ge_u32 %[cnt], 6\n"                     \
  "s_cbranch_scc1 L%=_nesc" NP1 "\n"             \
  "s_cmp_eq_u32 " MB_MAG ", 1\n"                 \
  "s_cselect_b32 %[t1], 8, 16\n"                 \
  "s_add_u32 %[off], %[offz], %[t1]\n"           \
  "s_add_u32 %[bf], %[bfz], %[t1]\n"             \
  "s_add_u32 %[sumabs], %[sumabs], " MB_MAG "\n" \
  MBP_LEVEL_##F(P, NP1, NP2, RB)
// A level in state a: the block's first, its sign deferred; state b follows.
#define MBP_LEVEL_a(P, NP1, NP2, RB)             \
  MBP_FIRST(NP1)
#define MBP_FIRST(NP)                            \
  "s_lshl_b32 %[t1], " MB_MAG ", 8\n"            \
  "s_or_b32 %[firstpm], %[t1], %[si]\n"          \
  "s_add_u32 %[firstsi], %[si], 1\n"             \
  "s_mov_b32 %[last], %[si]\n"                   \
  "s_add_u32 %[si], %[si], 1\n"                  \
  "s_cmp_lg_u32 %[si], %[stop]\n"                \
  "s_cbranch_scc1 L%=_tok" NP "b\n"              \
  "s_branch L%=_tstop" NP "b\n"
// A later level in state b: its explicit sign, then its entry.
#define MBP_LEVEL_b(P, NP1, NP2, RB)             \
  MBP_SIGN("ts" P, RB)                           \
  MBP_LATER(P, NP2)
#define MBP_LATER(TAG, NP)                       \
  MB_EMIT_AT("e1" TAG, "%[si]")                  \
  "s_mov_b32 %[last], %[si]\n"                   \
  "s_add_u32 %[si], %[si], 1\n"                  \
  "s_cmp_lg_u32 %[si], %[stop]\n"                \
  "s_cbranch_scc1 L%=_tok" NP "b\n"              \
  "s_branch L%=_tstop" NP "b\n"
// An escaped level (magnitude in MB_MAG) at phase 0, in either state.
#define MBP_ESCLEVEL                             \
  "L%=_tlev0:\n"                                 \
  "s_add_u32 %[sumabs], %[sumabs], " MB_MAG "\n" \
  "s_cmp_eq_u32 %[firstsi], 0\n"                 \
  "s_cbranch_scc1 L%=_tfirstg\n"                 \
  MBP_SIGN("tsg", "%[x0]")                       \
  MBP_LATER("g", "1")                            \
  "L%=_tfirstg:\n"                               \
  MBP_FIRST("0")
// si reached stop at phase P in state F: the end of the block (back to phase
// 0, then the end-of-block code) or a band boundary (bandOf thresholds 1, 3,
// 9, 21, model.go:165-180): the band's lane group alternates between offsets
// 0 and 24 and its vector advances every second band; the previous class
// carries over.
#define MBP_TSTOP(P, F)                          \
  "L%=_tstop" P F ":\n"                          \
  "s_cmp_eq_u32 %[si], %[total]\n"               \
  "s_cbranch_scc1 L%=_nend" P "\n"               \
  "s_sub_u32 %[t1], %[off], %[offz]\n"           \
  "s_xor_b32 %[offz], %[offz], 24\n"             \
  "s_xor_b32 %[bfz], %[bfz], 24\n"               \
  "s_add_u32 %[off], %[offz], %[t1]\n"           \
  "s_add_u32 %[bf], %[bfz], %[t1]\n"             \
  "s_cmp_eq_u32 %[si], 1\n"                      \
  "s_cselect_b32 %[stop], 3, %[stop]\n"          \
  "s_cmp_eq_u32 %[si], 9\n"                      \
  "s_cselect_b32 %[stop], 21, %[stop]\n"         \
  "s_cmp_eq_u32 %[si], 3\n"                      \
  "s_cselect_b32 %[stop], 9, %[stop]\n"          \
  "s_cselect_b64 %[b], -1, 0\n"                  \
  "v_cndmask_b32_e64 %[vcur], %[vcur], %[vb], %[b]\n" \
  "s_cmp_eq_u32 %[si], 21\n"                     \
  "s_cselect_b32 %[stop], %[total], %[stop]\n"   \
  "s_cselect_b64 %[b], -1, 0\n"                  \
  "v_cndmask_b32_e64 %[vcur], %[vcur], %[vc], %[b]\n" \
  "s_min_u32 %[stop], %[stop], %[total]\n"       \
  "s_branch L%=_tok" P F "\n"
// Back to phase 0 from phase 1, 2, 3 (t1 as the spare register).
#define MBP_NORM1                                \
  "s_mov_b32 %[t1], %[x0]\n"                     \
  "s_mov_b32 %[x0], %[x1]\n"                     \
  "s_mov_b32 %[x1], %[x2]\n"                     \
  "s_mov_b32 %[x2], %[x3]\n"                     \
  "s_mov_b32 %[x3], %[t1]\n"
#define MBP_NORM2                                \
  "s_mov_b32 %[t1], %[x0]\n"                     \
  "s_mov_b32 %[x0], %[x2]\n"                     \
  "s_mov_b32 %[x2], %[t1]\n"                     \
  "s_mov_b32 %[t1], %[x1]\n"                     \
  "s_mov_b32 %[x1], %[x3]\n"                     \
  "s_mov_b32 %[x3], %[t1]\n"
#define MBP_NORM3                                \
  "s_mov_b32 %[t1], %[x3]\n"                     \
  "s_mov_b32 %[x3], %[x2]\n"                     \
  "s_mov_b32 %[x2], %[x1]\n"                     \
  "s_mov_b32 %[x1], %[x0]\n"                     \
  "s_mov_b32 %[x0], %[t1]\n"
#define MBP_TOKLOOP                              \
  MBP_TOK("0", "1", "%[x0]", a)                  \
  MBP_TOK("1", "2", "%[x1]", a)                  \
  MBP_TOK("2", "3", "%[x2]", a)                  \
  MBP_TOK("3", "0", "%[x3]", a)                  \
  "s_branch L%=_tok0a\n"                         \
  MBP_TOK("0", "1", "%[x0]", b)                  \
  MBP_TOK("1", "2", "%[x1]", b)                  \
  MBP_TOK("2", "3", "%[x2]", b)                  \
  MBP_TOK("3", "0", "%[x3]", b)                  \
  "s_branch L%=_tok0b\n"                         \
  MBP_NZ("0", "1", "2", "%[x0]", "%[x1]", a)     \
  MBP_NZ("1", "2", "3", "%[x1]", "%[x2]", a)     \
  MBP_NZ("2", "3", "0", "%[x2]", "%[x3]", a)     \
  MBP_NZ("3", "0", "1", "%[x3]", "%[x0]", a)     \
  MBP_NZ("0", "1", "2", "%[x0]", "%[x1]", b)     \
  MBP_NZ("1", "2", "3", "%[x1]", "%[x2]", b)     \
  MBP_NZ("2", "3", "0", "%[x2]", "%[x3]", b)     \
  MBP_NZ("3", "0", "1", "%[x3]", "%[x0]", b)     \
  MBP_TSTOP("0", "a") MBP_TSTOP("1", "a") MBP_TSTOP("2", "a") MBP_TSTOP("3", "a") \
  MBP_TSTOP("0", "b") MBP_TSTOP("1", "b") MBP_TSTOP("2", "b") MBP_TSTOP("3", "b") \
  MBP_ESCLEVEL                                   \
  "L%=_nesc1:\n" MBP_NORM1 "s_branch L%=_nesc0\n" \
  "L%=_nesc2:\n" MBP_NORM2 "s_branch L%=_nesc0\n" \
  "L%=_nesc3:\n" MBP_NORM3 "s_branch L%=_nesc0\n" \
  "L%=_nend0:\n" "s_branch L%=_tend\n"           \
  "L%=_nend1:\n" MBP_NORM1 "s_branch L%=_tend\n" \
  "L%=_nend2:\n" MBP_NORM2 "s_branch L%=_tend\n" \
  "L%=_nend3:\n" MBP_NORM3 "s_branch L%=_tend\n"
#define MBP_TOKLOOP_OOL                          \
  MBP_RN_OOL("tz0a", "underrun", "%[x0]") MBP_RN_OOL("tn0a", "underrun", "%[x0]") \
  MBP_RN_OOL("tz1a", "underrun", "%[x1]") MBP_RN_OOL("tn1a", "underrun", "%[x1]") \
  MBP_RN_OOL("tz2a", "underrun", "%[x2]") MBP_RN_OOL("tn2a", "underrun", "%[x2]") \
  MBP_RN_OOL("tz3a", "underrun", "%[x3]") MBP_RN_OOL("tn3a", "underrun", "%[x3]") \
  MBP_RN_OOL("tz0b", "underrun", "%[x0]") MBP_RN_OOL("tn0b", "underrun", "%[x0]") \
  MBP_RN_OOL("tz1b", "underrun", "%[x1]") MBP_RN_OOL("tn1b", "underrun", "%[x1]") \
  MBP_RN_OOL("tz2b", "underrun", "%[x2]") MBP_RN_OOL("tn2b", "underrun", "%[x2]") \
  MBP_RN_OOL("tz3b", "underrun", "%[x3]") MBP_RN_OOL("tn3b", "underrun", "%[x3]") \
  MBP_RN_OOL("ts0", "underrun", "%[x1]") MB_EMIT_OOL("e10")                       \
  MBP_RN_OOL("ts1", "underrun", "%[x2]") MB_EMIT_OOL("e11")                       \
  MBP_RN_OOL("ts2", "underrun", "%[x3]") MB_EMIT_OOL("e12")                       \
  MBP_RN_OOL("ts3", "underrun", "%[x0]") MB_EMIT_OOL("e13")                       \
  MBP_RN_OOL("tsg", "underrun", "%[x0]") MB_EMIT_OOL("e1g")

// A block's first token (band 0, previous class 0, phase 0; the block's
// vectors in va, vb, vc): the zero token and the end of block (symbol 6: the
// ballot's field is 0x7f) are tried together.  A zero token goes to _tbz,
// any other token but the end of block to _tbnz; the end of block (a block
// with no level) falls through, decoded and rotated as MB_DEC leaves a
// symbol.
#define MB_FIRST_TOKEN(tag)                      \
  "s_and_b32 %[slot], %[x0], 0xfff\n"            \
  "v_cmp_ge_u32_sdwa %[b], %[slot], %[va] src0_sel:DWORD src1_sel:WORD_1\n" \
  "v_readlane_b32 %[p], %[va], 0\n"              \
  "v_readlane_b32 %[t3], %[va], 6\n"             \
  "v_mov_b32 %[vcur], %[va]\n"                   \
  "s_lshr_b32 %[x0], %[x0], 12\n"                \
  "s_bfe_u64 %[b], %[b], 0x80000\n"              \
  "s_mov_b32 %[sumabs], 0\n"                     \
  "s_mov_b32 %[firstsi], 0\n"                    \
  "s_cmp_eq_u64 %[b], 1\n"                       \
  "s_cbranch_scc1 L%=_tbz\n"                     \
  "s_cmp_eq_u64 %[b], 0x7f\n"                    \
  "s_cbranch_scc0 L%=_tbnz\n"                    \
  "s_and_b32 %[t2], %[t3], 0xffff\n"             \
  "s_lshr_b32 %[t3], %[t3], 16\n"                \
  "s_mul_i32 %[xn], %[t2], %[x0]\n"              \
  "s_sub_u32 %[t2], %[slot], %[t3]\n"            \
  "s_add_u32 %[xn], %[xn], %[t2]\n"              \
  MB_RN(tag)

// A coded macroblock's fields before its grammar: intra kind, nothing
// written, zero vectors and modes, every plane coded.
#define MB_INIT                                  \
  "s_mov_b32 %[r0], 2\n"                         \
  "s_mov_b32 %[r1], 0\n"                         \
  "s_mov_b32 %[wmask], 0\n"                      \
  "s_mov_b32 %[mv0], 0\n"                        \
  "s_mov_b32 %[mv1], 0\n"                        \
  "s_mov_b32 %[mv2], 0\n"                        \
  "s_mov_b32 %[mv3], 0\n"                        \
  "s_mov_b32 %[mvcur], 0\n"                      \
  "s_mov_b32 %[lm0], 0\n"                        \
  "s_mov_b32 %[lm1], 0\n"                        \
  "s_mov_b32 %[cbpf], 7\n"

// The records of a page of macroblocks are staged transposed: dword k of the
// macroblock in lane l sits in lane l of vector vr<k>.  A page starts as the
// record of a skip macroblock (kind 0, every block written, no coefficients,
// zero vectors and modes), so a skip macroblock writes only its coefficient
// offset and vector.
#define MB_REC_TEMPLATE                          \
  "v_mov_b32 %[vr0], 0\n"                        \
  "v_mov_b32 %[vr1], 0\n"                        \
  "v_mov_b32 %[vr2], 0x3ffff\n"                  \
  "v_mov_b32 %[vr4], 0\n"                        \
  "v_mov_b32 %[vr6], 0\n"                        \
  "v_mov_b32 %[vr7], 0\n"                        \
  "v_mov_b32 %[vr8], 0\n"                        \
  "v_mov_b32 %[vr9], 0\n"                        \
  "v_mov_b32 %[vr10], 0\n"                       \
  "v_mov_b32 %[vr11], 0\n"                       \
  "v_mov_b32 %[vmv], 0\n"

// decodeTile (decode_intra.go:121-135) as one asm statement: macroblock
// modes, MV prediction (mv.go:21-52), the inter header (decode_inter.go:18-62:
// reference, partition, MV class values with raw bits, coded-block pattern),
// transform size and intra modes (decode_intra.go:137-167), every token of
// every block with escapes, explicit and hidden signs (residual.go:79-126),
// coefficient entries, macroblock records and the MV rows.  All decoder state
// is uniform and lives in SGPRs; vector registers hold the symbol tables, the
// byte window, the coefficient staging register (64 entries), the staged
// records of a page and the MV rows.
//
// Macroblocks are walked in pages: up to 64 consecutive columns of a row, one
// lane each.  Per page the MVs of the row above (LDS) are loaded into two
// vectors, B (above) and C (above-right, or above-left at the tile's right
// edge, or above itself in a one-column tile); in the tile's first row B and
// C are the int16 maximum and minimum.  predictMV is then the packed int16
// median of (left, B, C) in lane `column - page base` for every macroblock:
// zero for a missing left, the left vector alone in the first row, the above
// vector alone in a one-column tile (mv.go:21-52 case by case).  A page's
// records and MVs go out when it ends.
//
// It returns at the end of the tile (kMbDone) or at a decode error
// (kMbUnused, kMbUnderrun) with the failing macroblock's fields decoded so
// far, which the caller turns into its record exactly as the generic path
// does; the records of the page before it are already stored.
// Precondition: the sign context has a table (signs from two SGPRs).
__device__ __forceinline__ uint32_t mb_loop(MbLoop& S, const Tables& T, uint32_t vlane4, uint32_t x0t, uint32_t x1t,
                                            uint32_t y0t, uint32_t y1t, uint32_t mbc, uint32_t interf, uint32_t tcb,
                                            uint32_t f0, uint32_t f1, uint32_t hi2, uint32_t wlo, const uint8_t* tb,
                                            uint32_t* dst, MBRec* recs) {
  uint32_t ev, slot, t1, t2, t3, p, xn, cnt, sym, val, k2;
  uint32_t k, nb, total, blk8, cbpf, si, stop, off, offz, bf, bfz, sumabs, firstsi, firstpm, last, m0s, ln, pend;
  uint64_t b;
  uint32_t vcur, va, vb, vc, vaddr, vt;
  uint32_t vr0, vr1, vr2, vr3, vr4, vr5, vr6, vr7, vr8, vr9, vr10, vr11, vmv, vB, vC;
  asm volatile(
      "s_nop 4\n"
      "s_mov_b32 %[m0s], m0\n"
      // the tile's first row: no row above
      "v_mov_b32 %[vB], 0x7fff7fff\n"
      "v_mov_b32 %[vC], 0x80008000\n"
      "s_mov_b32 %[leftmv], 0\n"

      // ------------------------------------------------------------- page
      "L%=_page:\n"
      "s_add_u32 %[pend], %[mbx], 64\n"
      "s_min_u32 %[pend], %[pend], %[x1t]\n"
      "s_mov_b32 %[ln], 0\n"
      MB_REC_TEMPLATE

      // ---------------------------------------------------------------- MB
      "L%=_mb:\n"
      "s_add_u32 %[cstart], %[cbase], %[cn]\n"
      "s_cmp_eq_u32 %[interf], 0\n"
      "s_cbranch_scc1 L%=_kmb\n"
      // inter frame: the mode (decode_inter.go:9-16) and the MV predictor in
      // one VALU group; symbol 0 (skip) is the straight line
      "s_and_b32 %[slot], %[x0], 0xfff\n"
      "v_cmp_ge_u32_sdwa %[b], %[slot], %[vm0] src0_sel:DWORD src1_sel:WORD_1\n"
      "v_readlane_b32 %[p], %[vm0], 16\n"
      "v_pk_min_i16 %[vaddr], %[leftmv], %[vB]\n"
      "v_pk_max_i16 %[vt], %[leftmv], %[vB]\n"
      "v_pk_min_i16 %[vt], %[vt], %[vC]\n"
      "v_pk_max_i16 %[vaddr], %[vaddr], %[vt]\n"
      "v_readlane_b32 " MB_PRED ", %[vaddr], %[ln]\n"
      "s_lshr_b32 %[xn], %[x0], 12\n"
      "s_bfe_u64 %[b], %[b], 0x100010\n"
      "s_cmp_eq_u64 %[b], 1\n"
      "s_cbranch_scc0 L%=_mdnz\n"
      "s_mul_i32 %[xn], %[p], %[xn]\n"
      "s_add_u32 %[xn], %[xn], %[slot]\n"
      MB_RN("md")
      // skip macroblock: the predicted vector, every block written
      "s_mov_b32 m0, %[ln]\n"
      "s_add_u32 %[t1], %[tcb], %[cstart]\n"
      "v_writelane_b32 %[vr3], %[t1], m0\n"
      "v_writelane_b32 %[vr5], " MB_PRED ", m0\n"
      "v_writelane_b32 %[vmv], " MB_PRED ", m0\n"
      "s_mov_b32 %[leftmv], " MB_PRED "\n"
      "L%=_mbnext:\n"
      "s_add_u32 %[done], %[done], 1\n"
      "s_add_u32 %[mbx], %[mbx], 1\n"
      "s_add_u32 %[ln], %[ln], 1\n"
      "s_cmp_lt_u32 %[mbx], %[pend]\n"
      "s_cbranch_scc1 L%=_mb\n"
      // end of a page: its records and MVs out
      "L%=_pageend:\n"
      "s_mov_b32 %[t3], 0\n"
      "s_branch L%=_flush\n"
      "L%=_pagedone:\n"
      "s_cmp_lt_u32 %[mbx], %[x1t]\n"
      "s_cbranch_scc1 L%=_abv\n"
      // end of a row: the rows swap
      "s_mov_b32 %[t1], %[prowa]\n"
      "s_mov_b32 %[prowa], %[crowa]\n"
      "s_mov_b32 %[crowa], %[t1]\n"
      "s_mov_b32 %[mbx], %[x0t]\n"
      "s_mov_b32 %[leftmv], 0\n"
      "s_add_u32 %[mby], %[mby], 1\n"
      "s_cmp_lt_u32 %[mby], %[y1t]\n"
      "s_cbranch_scc0 L%=_tiledone\n"
      // the row above at the page's columns: B = above, C = above-right, or
      // above-left past the tile's last column (above itself when the tile
      // is one column wide)
      "L%=_abv:\n"
      "s_cmp_eq_u32 %[mby], %[y0t]\n"
      "s_cbranch_scc1 L%=_page\n"
      "s_sub_u32 %[t1], %[mbx], %[x0t]\n"
      "s_lshl_b32 %[t1], %[t1], 2\n"
      "s_add_u32 %[t1], %[t1], %[prowa]\n"
      "s_sub_u32 %[t2], %[x1t], %[mbx]\n"
      "s_lshl_b32 %[t2], %[t2], 2\n"
      "s_add_u32 %[t2], %[t2], -4\n"
      "v_min_u32 %[vaddr], %[t2], %[vlane4]\n"
      "v_add_u32 %[vt], 4, %[vaddr]\n"
      "v_cmp_ge_u32_e32 vcc, %[t2], %[vt]\n"
      "v_add_u32 %[vaddr], %[t1], %[vaddr]\n"
      "v_add_u32 %[vt], %[t1], %[vt]\n"
      "v_add_u32 %[vr0], -4, %[vaddr]\n"
      "v_max_u32 %[vr0], %[prowa], %[vr0]\n"
      "v_cndmask_b32_e32 %[vt], %[vr0], %[vt], vcc\n"
      "ds_read_b32 %[vB], %[vaddr]\n"
      "ds_read_b32 %[vC], %[vt]\n"
      "s_waitcnt lgkmcnt(0)\n"
      "s_branch L%=_page\n"

      // ---- the staged records of the page (lanes [0, ln)) to recs; then
      // (t3 = 0) the page's MVs into the current row, or (t3 = 1) the exit
      "L%=_flush:\n"
      "s_cmp_eq_u32 %[ln], 0\n"
      "s_cbranch_scc1 L%=_flushed\n"
      "s_bfm_b64 %[b], %[ln], 0\n"
      "s_cmp_eq_u32 %[ln], 64\n"
      "s_cselect_b64 %[b], -1, %[b]\n"
      "s_sub_u32 %[t1], %[mbx], %[ln]\n"
      "s_mul_i32 %[t2], %[mby], %[mbc]\n"
      "s_add_u32 %[t2], %[t2], %[t1]\n"
      "s_mul_i32 %[t2], %[t2], 48\n"
      "v_mul_u32_u24 %[vaddr], 12, %[vlane4]\n"
      "v_add_u32 %[vaddr], %[t2], %[vaddr]\n"
      "s_mov_b64 exec, %[b]\n"
      "global_store_dword %[vaddr], %[vr0], %[recs]\n"
      "global_store_dword %[vaddr], %[vr1], %[recs] offset:4\n"
      "global_store_dword %[vaddr], %[vr2], %[recs] offset:8\n"
      "global_store_dword %[vaddr], %[vr3], %[recs] offset:12\n"
      "global_store_dword %[vaddr], %[vr4], %[recs] offset:16\n"
      "global_store_dword %[vaddr], %[vr5], %[recs] offset:20\n"
      "global_store_dword %[vaddr], %[vr6], %[recs] offset:24\n"
      "global_store_dword %[vaddr], %[vr7], %[recs] offset:28\n"
      "global_store_dword %[vaddr], %[vr8], %[recs] offset:32\n"
      "global_store_dword %[vaddr], %[vr9], %[recs] offset:36\n"
      "global_store_dword %[vaddr], %[vr10], %[recs] offset:40\n"
      "global_store_dword %[vaddr], %[vr11], %[recs] offset:44\n"
      "s_cmp_eq_u32 %[t3], 0\n"
      "s_cbranch_scc0 L%=_flushx\n"
      "s_sub_u32 %[t1], %[t1], %[x0t]\n"
      "s_lshl_b32 %[t1], %[t1], 2\n"
      "s_add_u32 %[t1], %[t1], %[crowa]\n"
      "v_add_u32 %[vt], %[t1], %[vlane4]\n"
      "ds_write_b32 %[vt], %[vmv]\n"
      "s_mov_b64 exec, -1\n"
      "s_branch L%=_pagedone\n"
      "L%=_flushx:\n"
      "s_mov_b64 exec, -1\n"
      "s_branch L%=_exit\n"
      "L%=_flushed:\n"
      "s_cmp_eq_u32 %[t3], 0\n"
      "s_cbranch_scc1 L%=_pagedone\n"
      "s_branch L%=_exit\n"
      "L%=_tiledone:\n"
      "s_mov_b32 %[ev], 0\n"
      "s_branch L%=_exit\n"

      // ---- a coded macroblock of an inter frame: the mode's other symbols
      "L%=_mdnz:\n"
      "s_bcnt1_i32_b64 %[sym], %[b]\n"
      "s_cmp_eq_u32 %[sym], 0\n"
      "s_cbranch_scc1 L%=_mode_unused\n"
      "s_add_u32 %[t1], %[sym], 15\n"
      "v_readlane_b32 %[p], %[vm0], %[t1]\n"
      "s_add_u32 %[sym], %[sym], -1\n"
      "s_and_b32 %[t2], %[p], 0xffff\n"
      "s_lshr_b32 %[p], %[p], 16\n"
      "s_mul_i32 %[xn], %[t2], %[xn]\n"
      "s_sub_u32 %[t2], %[slot], %[p]\n"
      "s_add_u32 %[xn], %[xn], %[t2]\n"
      MB_RN("md2")
      MB_INIT
      "s_cmp_eq_u32 %[sym], 2\n"
      "s_cbranch_scc1 L%=_tx\n"
      // ---- inter macroblock: reference, partition, vectors, pattern
      "s_mov_b32 %[r0], 1\n"
      MB_DEC("%[vm1]", "32", "31", "0x100020", "%[sym]", "rf", "unused")
      "s_mov_b32 %[r1], %[sym]\n"
      // golden: the vectors are coded against zero
      "s_cmp_eq_u32 %[sym], 1\n"
      "s_cselect_b32 " MB_PRED ", 0, " MB_PRED "\n"
      MB_DEC("%[vm1]", "48", "47", "0x100030", "%[sym]", "pt", "unused")
      "s_lshl_b32 %[t1], %[sym], 8\n"
      "s_or_b32 %[r0], %[r0], %[t1]\n"
      "s_lshl_b32 %[t1], %[sym], 2\n"
      "s_lshr_b32 " MB_NPART ", 0x4221, %[t1]\n"
      "s_and_b32 " MB_NPART ", " MB_NPART ", 15\n"
      "s_mov_b32 " MB_PI ", 0\n"
      "L%=_mvl:\n"
      "s_cmp_eq_u32 " MB_PI ", 0\n"
      "s_cselect_b32 %[cnt], " MB_PRED ", %[mv0]\n"
      // x component: base.x + zigzag(class value), clamped to +-4096
      MB_CLASS("%[vm2]", "16", "15", "0x100010", "cx")
      "s_sext_i32_i16 %[t1], %[cnt]\n"
      "s_lshr_b32 %[t2], %[sym], 1\n"
      "s_and_b32 %[sym], %[sym], 1\n"
      "s_sub_u32 %[sym], 0, %[sym]\n"
      "s_xor_b32 %[t2], %[t2], %[sym]\n"
      "s_add_u32 %[t1], %[t1], %[t2]\n"
      "s_max_i32 %[t1], %[t1], -4096\n"
      "s_min_i32 %[t1], %[t1], 4096\n"
      "s_and_b32 " MB_MVT ", %[t1], 0xffff\n"
      MB_CLASS("%[vm2]", "32", "31", "0x100020", "cy")
      "s_ashr_i32 %[t1], %[cnt], 16\n"
      "s_lshr_b32 %[t2], %[sym], 1\n"
      "s_and_b32 %[sym], %[sym], 1\n"
      "s_sub_u32 %[sym], 0, %[sym]\n"
      "s_xor_b32 %[t2], %[t2], %[sym]\n"
      "s_add_u32 %[t1], %[t1], %[t2]\n"
      "s_max_i32 %[t1], %[t1], -4096\n"
      "s_min_i32 %[t1], %[t1], 4096\n"
      "s_lshl_b32 %[t1], %[t1], 16\n"
      "s_or_b32 " MB_MVT ", " MB_MVT ", %[t1]\n"
      "s_cmp_eq_u32 " MB_PI ", 0\n"
      "s_cselect_b32 %[mv0], " MB_MVT ", %[mv0]\n"
      "s_cmp_eq_u32 " MB_PI ", 1\n"
      "s_cselect_b32 %[mv1], " MB_MVT ", %[mv1]\n"
      "s_cmp_eq_u32 " MB_PI ", 2\n"
      "s_cselect_b32 %[mv2], " MB_MVT ", %[mv2]\n"
      "s_cmp_eq_u32 " MB_PI ", 3\n"
      "s_cselect_b32 %[mv3], " MB_MVT ", %[mv3]\n"
      "s_add_u32 " MB_PI ", " MB_PI ", 1\n"
      "s_cmp_lt_u32 " MB_PI ", " MB_NPART "\n"
      "s_cbranch_scc1 L%=_mvl\n"
      "s_cmp_eq_u32 %[r1], 1\n"
      "s_cselect_b32 %[mvcur], 0, %[mv0]\n"
      MB_DEC("%[vm0]", "48", "47", "0x100030", "%[cbpf]", "cb", "unused")
      "s_lshl_b32 %[t1], %[cbpf], 16\n"
      "s_or_b32 %[r0], %[r0], %[t1]\n"
      "s_mov_b32 %[nb], 0\n"
      "s_bitcmp1_b32 %[cbpf], 0\n"
      "s_cbranch_scc1 L%=_tx\n"
      "s_mov_b32 %[wmask], 0xffff\n"
      "s_branch L%=_blocks\n"

      // ---- key frame macroblock (always intra)
      "L%=_kmb:\n"
      MB_INIT

      // ---- transform size; the luma blocks' vectors and scan
      "L%=_tx:\n"
      MB_DEC("%[vm0]", "32", "31", "0x100020", "%[sym]", "tx", "unused")
      "s_lshl_b32 %[t1], %[sym], 24\n"
      "s_or_b32 %[r0], %[r0], %[t1]\n"
      // 16x16 (the most common size) on the straight line, 4x4 and 8x8
      // out of line
      "s_cmp_lt_u32 %[sym], 2\n"
      "s_cbranch_scc1 L%=_tx01\n"
      "s_mov_b32 %[nb], 1\n"
      "s_movk_i32 %[total], 0x100\n"
      "v_mov_b32 %[va], %[t20]\n"
      "v_mov_b32 %[vb], %[t21]\n"
      "v_mov_b32 %[vc], %[t22]\n"

      // ---- blocks: luma 0..nb-1 (intra: each with its mode), then Cb, Cr
      // when coded
      "L%=_blocks:\n"
      "s_mov_b32 %[k], 0\n"
      "s_cmp_eq_u32 %[nb], 0\n"
      "s_cbranch_scc1 L%=_lumadone\n"
      "s_mov_b32 %[blk8], 0\n"
      "s_bitcmp1_b32 %[r0], 1\n"
      "s_cbranch_scc0 L%=_tb\n"
      // intra luma mode of block k (4 bits in decode order)
      "L%=_lmdec:\n"
      MB_DEC("%[vm1]", "0", "-1", "0x100000", "%[sym]", "lm", "unused")
      "s_and_b32 %[t1], %[k], 7\n"
      "s_lshl_b32 %[t1], %[t1], 2\n"
      "s_lshl_b32 %[t1], %[sym], %[t1]\n"
      "s_cmp_lt_u32 %[k], 8\n"
      "s_cselect_b32 %[t2], %[t1], 0\n"
      "s_or_b32 %[lm0], %[lm0], %[t2]\n"
      "s_xor_b32 %[t2], %[t2], %[t1]\n"
      "s_or_b32 %[lm1], %[lm1], %[t2]\n"

      // ------------------------------------------------------ token block
      // The block's first token; an end of block here is a block with no
      // level, which goes straight to the next block.
      "L%=_tb:\n"
      MB_FIRST_TOKEN("tbe")
      "s_branch L%=_tbdone\n"
      MBP_TOKLOOP
      // end of block or escape (magnitude 5 + getEscapeValue,
      // mv.go:124-128), at phase 0
      "L%=_nesc0:\n"
      "s_cmp_eq_u32 %[cnt], 7\n"
      "s_cbranch_scc1 L%=_teob\n"
      MB_CLASS("%[vm2]", "0", "-1", "0x100000", "es")
      "s_add_u32 " MB_MAG ", %[sym], 5\n"
      "s_add_u32 %[off], %[offz], 16\n"
      "s_add_u32 %[bf], %[bfz], 16\n"
      "s_branch L%=_tlev0\n"
      // end of block: the first level's sign, hidden in the level-sum parity
      // when last - first >= 4, else explicit
      "L%=_teob:\n"
      "s_add_u32 %[nsym], %[nsym], 1\n"
      "L%=_tend:\n"
      "s_add_u32 %[nsym], %[nsym], %[si]\n"
      "s_cmp_eq_u32 %[firstsi], 0\n"
      "s_cbranch_scc1 L%=_tbdone\n"
      "s_lshr_b32 " MB_MAG ", %[firstpm], 8\n"
      "s_add_u32 %[t1], %[firstsi], -1\n"
      "s_sub_u32 %[t1], %[last], %[t1]\n"
      "s_cmp_lt_i32 %[t1], 4\n"
      "s_cbranch_scc0 L%=_tpar\n"
      MB_SIGN("sd")
      "s_branch L%=_temit\n"
      "L%=_tpar:\n"
      "s_bitcmp1_b32 %[sumabs], 0\n"
      "s_cbranch_scc0 L%=_temit\n"
      "s_sub_u32 " MB_MAG ", 0, " MB_MAG "\n"
      "L%=_temit:\n"
      "s_and_b32 %[t2], %[firstpm], 0xff\n"
      MB_EMIT("e2")
      // After a block: the next luma block out of line; the chroma mode
      // after the last luma block out of line; the chroma blocks on the
      // straight line, each with its own copy of the first token, so that an
      // empty chroma block costs one taken branch.
      "L%=_tbdone:\n"
      "s_lshr_b32 %[t1], %[blk8], 8\n"
      "s_bitset1_b32 %[wmask], %[t1]\n"
      "s_add_u32 %[k], %[k], 1\n"
      "s_cmp_lt_u32 %[k], %[nb]\n"
      "s_cbranch_scc1 L%=_tbmore\n"
      "s_cmp_eq_u32 %[k], %[nb]\n"
      "s_cbranch_scc1 L%=_lumadone\n"
      // chroma block k - nb (0 Cb, 1 Cr, 2 done)
      "L%=_chnext:\n"
      "s_sub_u32 %[t1], %[k], %[nb]\n"
      "s_cmp_eq_u32 %[t1], 2\n"
      "s_cbranch_scc1 L%=_mbend\n"
      "s_add_u32 %[t3], %[t1], 1\n"
      "s_lshr_b32 %[t3], %[cbpf], %[t3]\n"
      "s_bitcmp1_b32 %[t3], 0\n"
      "s_cbranch_scc0 L%=_chskip\n"
      "s_add_u32 %[t1], %[t1], 16\n"
      "s_lshl_b32 %[blk8], %[t1], 8\n"
      "s_mov_b32 %[total], 64\n"
      "v_mov_b32 %[va], %[t30]\n"
      "v_mov_b32 %[vb], %[t31]\n"
      "v_mov_b32 %[vc], %[t32]\n"
      MB_FIRST_TOKEN("tbe2")
      "s_branch L%=_tbdone\n"
      // the next luma block: inter straight to its tokens, intra to its mode
      "L%=_tbmore:\n"
      "s_lshl_b32 %[blk8], %[k], 8\n"
      "s_bitcmp1_b32 %[r0], 1\n"
      "s_cbranch_scc0 L%=_tb\n"
      "s_branch L%=_lmdec\n"
      // luma complete: the chroma mode (intra) or the uncoded chroma planes
      // (inter)
      "L%=_lumadone:\n"
      "s_bitcmp1_b32 %[r0], 1\n"
      "s_cbranch_scc0 L%=_ic\n"
      MB_DEC("%[vm1]", "16", "15", "0x100010", "%[sym]", "cm", "unused")
      "s_lshl_b32 %[t1], %[sym], 8\n"
      "s_or_b32 %[r1], %[r1], %[t1]\n"
      "s_branch L%=_chnext\n"
      "L%=_ic:\n"
      "s_bitcmp1_b32 %[cbpf], 1\n"
      "s_cbranch_scc1 L%=_icr\n"
      "s_bitset1_b32 %[wmask], 16\n"
      "L%=_icr:\n"
      "s_bitcmp1_b32 %[cbpf], 2\n"
      "s_cbranch_scc1 L%=_chnext\n"
      "s_bitset1_b32 %[wmask], 17\n"
      "s_branch L%=_chnext\n"
      "L%=_chskip:\n"
      "s_add_u32 %[k], %[k], 1\n"
      "s_branch L%=_chnext\n"

      // ---- end of a coded macroblock: its record and MV, then the next
      // macroblock as at _mbnext
      "L%=_mbend:\n"
      "s_and_b32 %[t3], %[r0], 0xff\n"
      "s_cmp_eq_u32 %[t3], 2\n"
      "s_cselect_b32 %[t3], 1, 0\n"
      "s_add_u32 %[nintra], %[nintra], %[t3]\n"
      "s_mov_b32 m0, %[ln]\n"
      "s_add_u32 %[t1], %[tcb], %[cstart]\n"
      "s_add_u32 %[t3], %[cbase], %[cn]\n"
      "s_sub_u32 %[t3], %[t3], %[cstart]\n"
      "v_writelane_b32 %[vr0], %[r0], m0\n"
      "v_writelane_b32 %[vr1], %[r1], m0\n"
      "v_writelane_b32 %[vr2], %[wmask], m0\n"
      "v_writelane_b32 %[vr3], %[t1], m0\n"
      "v_writelane_b32 %[vr4], %[t3], m0\n"
      "v_writelane_b32 %[vr5], %[mv0], m0\n"
      "v_writelane_b32 %[vr6], %[mv1], m0\n"
      "v_writelane_b32 %[vr7], %[mv2], m0\n"
      "v_writelane_b32 %[vr8], %[mv3], m0\n"
      "v_writelane_b32 %[vr9], %[lm0], m0\n"
      "v_writelane_b32 %[vr10], %[lm1], m0\n"
      "v_writelane_b32 %[vmv], %[mvcur], m0\n"
      "s_mov_b32 %[leftmv], %[mvcur]\n"
      "s_add_u32 %[done], %[done], 1\n"
      "s_add_u32 %[mbx], %[mbx], 1\n"
      "s_add_u32 %[ln], %[ln], 1\n"
      "s_cmp_lt_u32 %[mbx], %[pend]\n"
      "s_cbranch_scc1 L%=_mb\n"
      "s_branch L%=_pageend\n"

      // ---- decode errors: the failing macroblock's fields stay as decoded
      "L%=_mode_unused:\n"
      "s_mov_b32 %[ev], 3\n"
      "s_branch L%=_modeerr\n"
      "L%=_mode_underrun:\n"
      "s_mov_b32 %[ev], 4\n"
      // a failed mode reads as a skip macroblock (Go reconstructs it)
      "L%=_modeerr:\n"
      MB_INIT
      "s_mov_b32 %[r0], 0\n"
      "s_mov_b32 %[mv0], " MB_PRED "\n"
      "s_mov_b32 %[mvcur], " MB_PRED "\n"
      "s_mov_b32 %[wmask], 0x3ffff\n"
      "s_branch L%=_errx\n"
      "L%=_unused:\n"
      "s_mov_b32 %[ev], 3\n"
      "s_branch L%=_errx\n"
      "L%=_underrun:\n"
      "s_mov_b32 %[ev], 4\n"
      "L%=_errx:\n"
      "s_mov_b32 %[t3], 1\n"
      "s_branch L%=_flush\n"

      // ---- transform sizes 4x4 and 8x8: the luma blocks' vectors and scan
      "L%=_tx01:\n"
      "s_cmp_eq_u32 %[sym], 0\n"
      "s_cbranch_scc0 L%=_tx1\n"
      "s_mov_b32 %[nb], 16\n"
      "s_mov_b32 %[total], 16\n"
      "v_mov_b32 %[va], %[t00]\n"
      "v_mov_b32 %[vb], %[t01]\n"
      "v_mov_b32 %[vc], %[t02]\n"
      "s_branch L%=_blocks\n"
      "L%=_tx1:\n"
      "s_mov_b32 %[nb], 4\n"
      "s_mov_b32 %[total], 64\n"
      "v_mov_b32 %[va], %[t10]\n"
      "v_mov_b32 %[vb], %[t11]\n"
      "v_mov_b32 %[vc], %[t12]\n"
      "s_branch L%=_blocks\n"

      // ---- a block's first token when it is not the end of block: a zero
      // token (band 0 to band 1, as _tstop1a would) or any other (as _tok0a
      // leaves it for _tnz0a)
      "L%=_tbz:\n"
      "s_mul_i32 %[x0], %[p], %[x0]\n"
      "s_add_u32 %[x0], %[x0], %[slot]\n"
      MBP_RN_NC("tbz", "%[x0]")
      "s_mov_b32 %[si], 1\n"
      "s_mov_b32 %[stop], 3\n"
      "s_movk_i32 %[off], 24\n"
      "s_movk_i32 %[offz], 24\n"
      "s_mov_b32 %[bf], 0x80018\n"
      "s_mov_b32 %[bfz], 0x80018\n"
      "s_branch L%=_tok1a\n"
      "L%=_tbnz:\n"
      "s_mov_b32 %[si], 0\n"
      "s_mov_b32 %[stop], 1\n"
      "s_mov_b32 %[off], 0\n"
      "s_mov_b32 %[offz], 0\n"
      "s_mov_b32 %[bf], 0x80000\n"
      "s_mov_b32 %[bfz], 0x80000\n"
      "s_branch L%=_tnz0a\n"

      // ---- out-of-line renormalisation and coefficient stores
      MB_RN_OOL("md", "mode_underrun")
      MB_RN_OOL("md2", "mode_underrun")
      MB_RN_OOL("rf", "underrun")
      MB_RN_OOL("pt", "underrun")
      MB_RN_OOL("cxd", "underrun")
      MB_RN_OOL("cxb", "underrun")
      MB_RN_OOL("cyd", "underrun")
      MB_RN_OOL("cyb", "underrun")
      MB_RN_OOL("cb", "underrun")
      MB_RN_OOL("tx", "underrun")
      MB_RN_OOL("cm", "underrun")
      MB_RN_OOL("lm", "underrun")
      MB_RN_OOL("esd", "underrun")
      MB_RN_OOL("esb", "underrun")
      MB_RN_OOL("sd", "underrun")
      MB_RN_OOL("tbe", "underrun")
      MB_RN_OOL("tbe2", "underrun")
      MBP_RN_OOL("tbz", "underrun", "%[x0]")
      MBP_TOKLOOP_OOL
      MB_EMIT_OOL("e2")

      "L%=_exit:\n"
      "s_mov_b32 m0, %[m0s]\n"
      "s_nop 4\n"
      : [x0] "+&s"(S.x0), [x1] "+&s"(S.x1), [x2] "+&s"(S.x2), [x3] "+&s"(S.x3), [pos] "+&s"(S.pos),
        [nsym] "+&s"(S.nsym), [wbase] "+&s"(S.wbase), [lim] "+&s"(S.lim), [cn] "+&s"(S.cn), [cbase] "+&s"(S.cbase),
        [mbx] "+&s"(S.mbx), [mby] "+&s"(S.mby), [done] "+&s"(S.done), [nintra] "+&s"(S.nintra),
        [leftmv] "+&s"(S.leftmv), [vwin] "+&v"(S.vwin), [vbuf] "+&v"(S.vbuf),
        [prowa] "+&s"(S.prowa), [crowa] "+&s"(S.crowa), [r0] "=&s"(S.r0), [r1] "=&s"(S.r1),
        [wmask] "=&s"(S.wmask), [cstart] "=&s"(S.cstart), [mv0] "=&s"(S.mv0), [mv1] "=&s"(S.mv1),
        [mv2] "=&s"(S.mv2), [mv3] "=&s"(S.mv3), [mvcur] "=&s"(S.mvcur), [lm0] "=&s"(S.lm0), [lm1] "=&s"(S.lm1),
        [ev] "=&s"(ev), [slot] "=&s"(slot), [t1] "=&s"(t1), [t2] "=&s"(t2), [t3] "=&s"(t3), [p] "=&s"(p),
        [xn] "=&s"(xn), [cnt] "=&s"(cnt), [sym] "=&s"(sym), [val] "=&s"(val), [k2] "=&s"(k2), [k] "=&s"(k),
        [nb] "=&s"(nb), [total] "=&s"(total), [blk8] "=&s"(blk8), [cbpf] "=&s"(cbpf), [si] "=&s"(si),
        [stop] "=&s"(stop), [off] "=&s"(off), [offz] "=&s"(offz), [bf] "=&s"(bf), [bfz] "=&s"(bfz),
        [sumabs] "=&s"(sumabs), [firstsi] "=&s"(firstsi), [firstpm] "=&s"(firstpm), [last] "=&s"(last),
        [m0s] "=&s"(m0s), [ln] "=&s"(ln), [pend] "=&s"(pend), [b] "=&s"(b), [vcur] "=&v"(vcur), [va] "=&v"(va),
        [vb] "=&v"(vb), [vc] "=&v"(vc), [vaddr] "=&v"(vaddr), [vt] "=&v"(vt), [vr0] "=&v"(vr0),
        [vr1] "=&v"(vr1), [vr2] "=&v"(vr2), [vr3] "=&v"(vr3), [vr4] "=&v"(vr4), [vr5] "=&v"(vr5),
        [vr6] "=&v"(vr6), [vr7] "=&v"(vr7), [vr8] "=&v"(vr8), [vr9] "=&v"(vr9), [vr10] "=&v"(vr10),
        [vr11] "=&v"(vr11), [vmv] "=&v"(vmv), [vB] "=&v"(vB), [vC] "=&v"(vC)
      : [vm0] "v"(T.misc0), [vm1] "v"(T.misc1), [vm2] "v"(T.misc2), [t00] "v"(T.t00), [t01] "v"(T.t01),
        [t02] "v"(T.t02), [t10] "v"(T.t10), [t11] "v"(T.t11), [t12] "v"(T.t12), [t20] "v"(T.t20),
        [t21] "v"(T.t21), [t22] "v"(T.t22), [t30] "v"(T.t30), [t31] "v"(T.t31), [t32] "v"(T.t32),
        [vlane4] "v"(vlane4), [x0t] "s"(sgpr(x0t)), [x1t] "s"(sgpr(x1t)), [y0t] "s"(sgpr(y0t)),
        [y1t] "s"(sgpr(y1t)), [mbc] "s"(sgpr(mbc)), [interf] "s"(sgpr(interf)), [tcb] "s"(sgpr(tcb)),
        [f0] "s"(sgpr(f0)), [f1] "s"(sgpr(f1)), [hi2] "s"(sgpr(hi2)), [wlo] "s"(sgpr(wlo)), [tb] "s"(sgpr(tb)),
        [dst] "s"(sgpr(dst)), [recs] "s"(sgpr(recs))
      : "scc", "vcc", "memory");
  return ev;
}

// Pointers read from a job table are generic pointers; the job table is read
// through the constant address space, and the compiler takes generic
// pointers loaded from constant memory to be global ones, so accesses
// through a job's buffers are global (not flat) and uniform loads can be
// scalar.
using ConstJob = __attribute__((address_space(4))) const FrameJob;
using ConstItem = __attribute__((address_space(4))) const WorkItem;
__device__ __forceinline__ ConstJob& job_of(const FrameJob* jobs, uint32_t i) {
  return ((ConstJob*)jobs)[i];
}
__device__ __forceinline__ Planes planes_of(const __attribute__((address_space(4))) Planes& p) {
  return Planes{p.y, p.cb, p.cr, p.pw, p.ph, p.cw, p.ch};
}
__device__ __forceinline__ WorkItem item_of(const WorkItem* items) {
  const ConstItem& it = ((ConstItem*)items)[blockIdx.x];
  return WorkItem{it.job, it.idx};
}

// Workgroups are dispatched to the 8 XCDs of the chip round robin
// (workgroup b to XCD b % 8), and each XCD has its own L2.  This bijection
// of [0, n) gives XCD x a contiguous range of the work items instead, so
// neighbouring macroblocks (which share reference lines) meet in one L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t b, uint32_t n) {
  const uint32_t q = n / kXcds, r = n % kXcds, x = b % kXcds, k = b / kXcds;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Marks the records of tile td from local macroblock `done` on as never
// reached (kind kKindUnreached), so that reconstruction needs no tile status.
__device__ void mark_unreached(MBRec* __restrict__ recs, uint32_t mbc, const TileDesc& td, uint32_t done, int lane) {
  const uint32_t tw = td.x1 - td.x0, total = tw * (td.y1 - td.y0);
  for (uint32_t i = done + static_cast<uint32_t>(lane); i < total; i += 64) {
    const uint32_t yy = i / tw, xx = i - yy * tw;
    *reinterpret_cast<uint32_t*>(recs + (td.y0 + yy) * mbc + td.x0 + xx) = kKindUnreached;
  }
}

// decodeTile (decode_intra.go:121-135) for tile t: every macroblock in raster
// order until the first error; one record per macroblock reached.
constexpr uint32_t kRecStage = 16;  // macroblock records staged in LDS per store burst

// base[off / 4] = v on the lanes where pred holds.  Written as one asm
// statement: the compiler treats a store's operand registers as busy until
// the store completes (vmcnt), so a compiler-visible store inside the
// macroblock loop makes the next write of those registers wait for memory.
// The hardware reads a dword store's operands at issue; the asm keeps the
// compiler from waiting.
__device__ __forceinline__ void store_dword_if(uint32_t* base, uint32_t off, uint32_t v, bool pred) {
  uint64_t m, save;
  asm volatile(
      "s_nop 4\n"
      "v_cmp_ne_u32_e64 %[m], %[p], 0\n"
      "s_and_saveexec_b64 %[save], %[m]\n"
      "global_store_dword %[off], %[v], %[base]\n"
      "s_mov_b64 exec, %[save]\n"
      : [m] "=&s"(m), [save] "=&s"(save)
      : [p] "v"(static_cast<uint32_t>(pred)), [off] "v"(off), [v] "v"(v), [base] "s"(sgpr(base))
      : "memory", "scc");
}

__device__ __forceinline__ void parse_tile(const uint8_t* __restrict__ pkt, const uint8_t* __restrict__ desc,
                                           MBRec* __restrict__ recs, uint32_t* __restrict__ coefs,
                                           TileStatus* __restrict__ status, const int t) {
  extern __shared__ uint32_t smem[];
  const FrameHdr* hdr = reinterpret_cast<const FrameHdr*>(desc);
  const TileDesc* tiles = reinterpret_cast<const TileDesc*>(desc + desc_tiles_off());
  const int lane = lane_id();
  const TileDesc td = tiles[t];

  // The frame's symbol tables, one entry per lane of each vector.
  using GWord = __attribute__((address_space(1))) const uint32_t;
  const GWord* vt = (const GWord*)(&hdr->vt[0][lane]);
  Tables T;
  T.misc0 = vt[0];
  T.misc1 = vt[64];
  T.misc2 = vt[128];
  T.t00 = vt[64 * 3]; T.t01 = vt[64 * 4]; T.t02 = vt[64 * 5];
  T.t10 = vt[64 * 6]; T.t11 = vt[64 * 7]; T.t12 = vt[64 * 8];
  T.t20 = vt[64 * 9]; T.t21 = vt[64 * 10]; T.t22 = vt[64 * 11];
  T.t30 = vt[64 * 12]; T.t31 = vt[64 * 13]; T.t32 = vt[64 * 14];
  // LDS: the macroblock records being staged, then two MV rows of the tile
  // width.  Records go to memory kRecStage at a time, every lane storing
  // dwords: a store per macroblock from one lane would have the next
  // macroblock's record registers wait for it.
  LdsU32* stage = (LdsU32*)smem;
  LdsU32* prow = (LdsU32*)(smem + kRecStage * sizeof(MBRec) / 4);
  LdsU32* crow = prow + hdr->max_tile_w;
  uint32_t nst = 0, first_mb = 0;
  auto flush = [&]() {
    const LdsU32* src = stage;
    uint32_t* dst = reinterpret_cast<uint32_t*>(recs + first_mb);
    const uint32_t nd = nst * (sizeof(MBRec) / 4);
    for (uint32_t i0 = 0; i0 < nd; i0 += 64) {
      const uint32_t i = i0 + static_cast<uint32_t>(lane);
      store_dword_if(dst + i0, 4u * static_cast<uint32_t>(lane), src[i < nd ? i : 0], i < nd);
    }
    nst = 0;
  };
  const int tw = td.x1 - td.x0;

  if (td.byte_len < 16) {  // ransDecoder.reset (rans.go:237-250)
    mark_unreached(recs, static_cast<uint32_t>(hdr->mbc), td, 0, lane);
    if (lane == 0) status[t] = TileStatus{kErrCorrupt | (kTileShort << 8), 0, 0, 0};
    return;
  }
  Rans R;
  R.win.init(pkt + td.byte_off, td.byte_len, lane);
  R.pos = td.byte_len;
  R.nsym = 0;
  R.err = 0;
  {
    uint32_t sv[4];
#pragma unroll
    for (int i = 3; i >= 0; i--) {  // pop32: big-endian from the top, state 3 first
      uint32_t a = R.win.byte(static_cast<int32_t>(--R.pos), lane);
      uint32_t b = R.win.byte(static_cast<int32_t>(--R.pos), lane);
      uint32_t c = R.win.byte(static_cast<int32_t>(--R.pos), lane);
      uint32_t e = R.win.byte(static_cast<int32_t>(--R.pos), lane);
      sv[i] = (a << 24) | (b << 16) | (c << 8) | e;
    }
    R.x0 = sv[0]; R.x1 = sv[1]; R.x2 = sv[2]; R.x3 = sv[3];
  }
  // The loop emits scan indices (reconstruction maps them to positions).
  const uint32_t zz4 = 0, zz8 = 0, zz16 = 0;
  // Let every start-up load land here (an empty asm that takes the loaded
  // registers), so no load into a register the macroblock loop uses is
  // outstanding there: a wait in the loop would also wait for every store
  // issued before it.
  asm volatile("" : "+v"(T.misc0), "+v"(T.misc1), "+v"(T.misc2));
  asm volatile("" : "+v"(T.t00), "+v"(T.t01), "+v"(T.t02), "+v"(T.t10), "+v"(T.t11), "+v"(T.t12));
  asm volatile("" : "+v"(T.t20), "+v"(T.t21), "+v"(T.t22), "+v"(T.t30), "+v"(T.t31), "+v"(T.t32));

  CoefOut out;
  out.dst = coefs + td.coef_base;
  out.base = 0; out.cn = 0; out.buf = 0;
  const int mbc = hdr->mbc;
  const bool inter_frame = hdr->inter != 0;
  uint32_t done = 0, nintra = 0;
  bool ok = true;
  for (int mby = td.y0; ok && mby < td.y1; mby++) {
    for (int mbx = td.x0; mbx < td.x1; mbx++) {
      if (nst == 0) first_mb = static_cast<uint32_t>(mby * mbc + mbx);
      MBState m{2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      const uint32_t coef_start = out.base + out.cn;
      bool coded = true;
      if (inter_frame) {
        // A failed mode decode reads as skip, which Go then reconstructs.
        const int mode = R.decode<kGMBMode, 16>(T.misc0, lane);
        if (mode < 0 || mode == kMBSkip) {
          const uint32_t pred = predict_mv(prow, crow, mbx - td.x0, tw, mby == td.y0);
          m.kind = 0; m.mv0 = pred; m.mvcur = pred; m.wmask = 0x3FFFFu;
          ok = mode >= 0;
          coded = false;
        } else if (mode == kMBInter) {
          m.kind = 1;
        }
      }
      if (coded) {
        const bool intra = m.kind == 2;
        const uint32_t pred = intra ? 0u : predict_mv(prow, crow, mbx - td.x0, tw, mby == td.y0);
        ok = parse_mb(R, out, m, T, zz4, zz8, zz16, intra, pred, lane);
      }
      if (m.kind == 2) nintra++;
      if (lane == 0) {
        crow[mbx - td.x0] = m.mvcur;
        MBRec r;
        r.kind = static_cast<uint8_t>(m.kind);
        r.part = static_cast<uint8_t>(m.part);
        r.cbp = static_cast<uint8_t>(m.cbp);
        r.tx = static_cast<uint8_t>(m.tx);
        r.refsel = static_cast<uint8_t>(m.refsel);
        r.cmode = static_cast<uint8_t>(m.cmode);
        r.pad0 = 0;
        r.pad1 = 0;
        r.wmask = m.wmask;
        r.coef_off = td.coef_base + coef_start;
        r.coef_count = out.base + out.cn - coef_start;
        r.mv[0][0] = mv_comp(m.mv0, 0); r.mv[0][1] = mv_comp(m.mv0, 1);
        r.mv[1][0] = mv_comp(m.mv1, 0); r.mv[1][1] = mv_comp(m.mv1, 1);
        r.mv[2][0] = mv_comp(m.mv2, 0); r.mv[2][1] = mv_comp(m.mv2, 1);
        r.mv[3][0] = mv_comp(m.mv3, 0); r.mv[3][1] = mv_comp(m.mv3, 1);
        r.lmode[0] = m.lmode0;
        r.lmode[1] = m.lmode1;
        r.pad2 = 0;
        const uint32_t* rw = reinterpret_cast<const uint32_t*>(&r);
        LdsU32* sw = stage + nst * (sizeof(MBRec) / 4);
#pragma unroll
        for (int k = 0; k < static_cast<int>(sizeof(MBRec) / 4); k++) sw[k] = rw[k];
      }
      done++;
      if (++nst == kRecStage) flush();
      if (!ok) break;
    }
    if (nst) flush();
    wave_sync();
    LdsU32* tmp = prow; prow = crow; crow = tmp;
  }
  if (static_cast<uint32_t>(lane) < out.cn) out.dst[out.base + lane] = out.buf;
  if (R.err) mark_unreached(recs, static_cast<uint32_t>(hdr->mbc), td, done, lane);
  if (lane == 0) status[t] = TileStatus{R.err, done, R.nsym, nintra};
}


// decodeTile through mb_loop (frames with a sign table); same outputs as
// parse_tile.
__device__ __forceinline__ void parse_tile_asm(const uint8_t* __restrict__ pkt, const uint8_t* __restrict__ desc,
                                           MBRec* __restrict__ recs, uint32_t* __restrict__ coefs,
                                           TileStatus* __restrict__ status, const int t) {
  extern __shared__ uint32_t smem[];
  const FrameHdr* hdr = reinterpret_cast<const FrameHdr*>(desc);
  const TileDesc* tiles = reinterpret_cast<const TileDesc*>(desc + desc_tiles_off());
  const int lane = lane_id();
  const TileDesc td = tiles[t];

  // The frame's symbol tables, one entry per lane of each vector.
  using GWord = __attribute__((address_space(1))) const uint32_t;
  const GWord* vt = (const GWord*)(&hdr->vt[0][lane]);
  Tables T;
  T.misc0 = vt[0];
  T.misc1 = vt[64];
  T.misc2 = vt[128];
  T.t00 = vt[64 * 3]; T.t01 = vt[64 * 4]; T.t02 = vt[64 * 5];
  T.t10 = vt[64 * 6]; T.t11 = vt[64 * 7]; T.t12 = vt[64 * 8];
  T.t20 = vt[64 * 9]; T.t21 = vt[64 * 10]; T.t22 = vt[64 * 11];
  T.t30 = vt[64 * 12]; T.t31 = vt[64 * 13]; T.t32 = vt[64 * 14];
  // The sign context's two frequencies (symbol 0 starts at 0).
  const uint32_t f0 = rd_lane(T.misc0, static_cast<int>(kGSign)) & 0xFFFFu;
  const uint32_t f1 = rd_lane(T.misc0, static_cast<int>(kGSign) + 1) & 0xFFFFu;
  const uint32_t vlane4 = 4u * static_cast<uint32_t>(lane);
  // LDS: two MV rows of the widest tile (the row above, the current row),
  // after the area the generic decoder stages records in.
  LdsU32* prow = (LdsU32*)(smem + kRecStage * sizeof(MBRec) / 4);
  LdsU32* crow = prow + hdr->max_tile_w;

  if (td.byte_len < 16) {  // ransDecoder.reset (rans.go:237-250)
    mark_unreached(recs, static_cast<uint32_t>(hdr->mbc), td, 0, lane);
    if (lane == 0) status[t] = TileStatus{kErrCorrupt | (kTileShort << 8), 0, 0, 0};
    return;
  }
  Rans R;
  R.win.init(pkt + td.byte_off, td.byte_len, lane);
  R.pos = td.byte_len;
  R.nsym = 0;
  R.err = 0;
  {
    uint32_t sv[4];
#pragma unroll
    for (int i = 3; i >= 0; i--) {  // pop32: big-endian from the top, state 3 first
      uint32_t a = R.win.byte(static_cast<int32_t>(--R.pos), lane);
      uint32_t b = R.win.byte(static_cast<int32_t>(--R.pos), lane);
      uint32_t c = R.win.byte(static_cast<int32_t>(--R.pos), lane);
      uint32_t e = R.win.byte(static_cast<int32_t>(--R.pos), lane);
      sv[i] = (a << 24) | (b << 16) | (c << 8) | e;
    }
    R.x0 = sv[0]; R.x1 = sv[1]; R.x2 = sv[2]; R.x3 = sv[3];
  }
  // Let every start-up load land here (an empty asm that takes the loaded
  // registers), so no load into a register the macroblock loop uses is
  // outstanding there: a wait in the loop would also wait for every store
  // issued before it.
  asm volatile("" : "+v"(T.misc0), "+v"(T.misc1), "+v"(T.misc2));
  asm volatile("" : "+v"(T.t00), "+v"(T.t01), "+v"(T.t02), "+v"(T.t10), "+v"(T.t11), "+v"(T.t12));
  asm volatile("" : "+v"(T.t20), "+v"(T.t21), "+v"(T.t22), "+v"(T.t30), "+v"(T.t31), "+v"(T.t32));

  uint32_t* const cdst = coefs + td.coef_base;
  const int mbc = hdr->mbc;
  const bool inter_frame = hdr->inter != 0;
  MbLoop S;
  S.x0 = sgpr(R.x0); S.x1 = sgpr(R.x1); S.x2 = sgpr(R.x2); S.x3 = sgpr(R.x3); S.pos = sgpr(R.pos);
  S.nsym = sgpr(R.nsym); S.wbase = sgpr(static_cast<uint32_t>(R.win.wbase));
  S.lim = sgpr(static_cast<uint32_t>(R.win.lim));
  S.cn = 0; S.cbase = 0;
  S.mbx = sgpr(static_cast<uint32_t>(td.x0)); S.mby = sgpr(static_cast<uint32_t>(td.y0));
  S.done = 0; S.nintra = 0; S.leftmv = 0;
  S.vwin = R.win.cur; S.vbuf = 0;
  S.prowa = sgpr(static_cast<uint32_t>(reinterpret_cast<uintptr_t>(prow)));
  S.crowa = sgpr(static_cast<uint32_t>(reinterpret_cast<uintptr_t>(crow)));
  const uint32_t hi2 = static_cast<uint32_t>(R.win.hi - R.win.lo - 4), wlo = static_cast<uint32_t>(R.win.lo);
  const uint32_t ev = mb_loop(S, T, vlane4, td.x0, td.x1, td.y0, td.y1, static_cast<uint32_t>(mbc),
                              inter_frame ? 1u : 0u, td.coef_base, f0, f1, hi2, wlo, R.win.tile + R.win.lo, cdst,
                              recs);
  uint32_t done = S.done, nintra = S.nintra;
  int32_t err = 0;
  if (ev != kMbDone) {
    err = static_cast<int32_t>(kErrCorrupt | ((ev == kMbUnused ? kTileUnused : kTileUnderrun) << 8));
    // the failing macroblock's record as decoded (the ones before it are stored)
    if ((S.r0 & 0xFFu) == 2) nintra++;
    if (lane == 0) {
      uint32_t* rw = reinterpret_cast<uint32_t*>(recs + S.mby * static_cast<uint32_t>(mbc) + S.mbx);
      rw[0] = S.r0; rw[1] = S.r1; rw[2] = S.wmask; rw[3] = td.coef_base + S.cstart;
      rw[4] = S.cbase + S.cn - S.cstart;
      rw[5] = S.mv0; rw[6] = S.mv1; rw[7] = S.mv2; rw[8] = S.mv3; rw[9] = S.lm0; rw[10] = S.lm1; rw[11] = 0;
    }
    done++;
  }
  if (static_cast<uint32_t>(lane) < S.cn) cdst[S.cbase + lane] = S.vbuf;
  if (err) mark_unreached(recs, static_cast<uint32_t>(hdr->mbc), td, done, lane);
  if (lane == 0) status[t] = TileStatus{err, done, S.nsym, nintra};
}

// One tile: its job's packet and descriptor, the entropy outputs.
// The band counters after a payload's residuals (key_bands).
__device__ __forceinline__ uint32_t* band_sync(int16_t* res, int mbc, int mbr) {
  return reinterpret_cast<uint32_t*>(res + static_cast<size_t>(mbc) * mbr * kResPerMB);
}

__device__ __forceinline__ void parse_item(const FrameJob* __restrict__ jobs, const WorkItem it) {
  ConstJob& job = job_of(jobs, it.job);
  const FrameHdr* hdr = reinterpret_cast<const FrameHdr*>(job.desc);
  // Bits 30-31 of the item: the wave's issue priority (the host gives the
  // biggest tiles, the longest serial chains, the highest).
  const uint32_t prio = it.idx >> 30;
  if (prio == 3)
    __builtin_amdgcn_s_setprio(3);
  else if (prio == 2)
    __builtin_amdgcn_s_setprio(2);
  else if (prio == 1)
    __builtin_amdgcn_s_setprio(1);
  const int t = static_cast<int>(it.idx & 0x3FFFFFFFu);
  {  // the tile's band counters for the intra launch
    const TileDesc td = reinterpret_cast<const TileDesc*>(job.desc + desc_tiles_off())[t];
    uint32_t* sync = band_sync(job.res, job.mbc, job.mbr);
    const int th = td.y1 - td.y0;
    for (int i0 = 0; i0 < th; i0 += 64) {  // (a uniform loop: the tile decoder's operands stay scalar)
      const int i = i0 + lane_id();
      if (i < th) sync[(td.y0 + i) * job.mbc + td.x0] = 0;
    }
  }
  if (hdr->vt[kSignVec][kSignLane] != kNoSym)
    parse_tile_asm(job.pkt, job.desc, job.recs, job.coefs, job.status, t);
  else
    parse_tile(job.pkt, job.desc, job.recs, job.coefs, job.status, t);
}

// The entropy launch: one wavefront per tile, item blockIdx.x (the host
// orders the items longest tile first: a tile's coded size predicts its
// serial chain, and wavefronts are dispatched in block order).
__global__ __launch_bounds__(64) void ebiv_parse(const FrameJob* __restrict__ jobs, const WorkItem* __restrict__ items) {
  parse_item(jobs, item_of(items));
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
// Returns the blocks holding a nonzero coefficient (bits 0-17) and, shifted
// by 32, those holding one away from the DC position.
__device__ uint64_t load_coefs(WaveScratch& ws, const uint32_t* __restrict__ coefs, const MBRec& r, int n, int qdc,
                               int qac, int lane) {
  for (int i = lane; i < 384; i += 64) ws.coef[i] = 0;
  wave_sync();
  uint32_t nzm = 0, acm = 0;
  uint32_t cnt = r.coef_count < 384u ? r.coef_count : 384u;
  const int zl = n == 4 ? 0 : (n == 8 ? 1 : 2);
  for (uint32_t i = lane; i < cnt; i += 64) {
    uint32_t e = coefs[r.coef_off + i];
    int si = e & 0xFF, blk = (e >> 8) & 31;
    int32_t level = static_cast<int32_t>(e) >> 13;
    int32_t v = level * (si == 0 ? qdc : qac);  // scan index 0 is the DC position
    int idx, pos;
    if (blk < 16) {
      int nn = n * n;
      if (si >= nn) si = nn - 1;
      pos = c_zz[zl][si];
      idx = blk * nn + pos;
      if (idx >= 256) continue;
    } else if (blk < 18) {
      si &= 63;
      pos = c_zz[1][si];
      idx = 256 + (blk - 16) * 64 + pos;
    } else {
      continue;
    }
    ws.coef[idx] = v;
    if (v != 0) {
      nzm |= 1u << blk;
      if (pos != 0) acm |= 1u << blk;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    nzm |= __shfl_xor(nzm, off);
    acm |= __shfl_xor(acm, off);
  }
  wave_sync();
  return static_cast<uint64_t>(nzm) | (static_cast<uint64_t>(acm) << 32);
}

// The DCT matrices of buildDCTMatrix (transform.go:15-40) for n = 4, 8, 16
// (c_dct rows 0, 1, 2) are staged into LDS by the wavefront that inverts a
// residual (a lane-indexed read of constant memory is a global load per
// element).
using LdsI32 = __attribute__((address_space(3))) int32_t;

// Separable inverse DCT (transform.go:129-194) of nblk n x n blocks at base:
// column pass into tmp, row pass back into coef, each with the
// (x + 4096) >> 13 rounding.  Zero coefficients add nothing, so this equals
// the reference's sparse forms bit for bit; in particular an all-zero block
// stays zero and a DC-only block is the constant
//   (c0 * ((c0 * dc + 4096) >> 13) + 4096) >> 13,  c0 = C[0][i] for every i
// (transform.go:141-160).  The reference accumulates in 64 bits; when every
// input is at most kIdct32Max in magnitude both passes' sums stay below 2^31
// (a column of the n = 16 matrix sums to at most 45,488 in magnitude, so the
// first pass is at most 5.56x the input and the second 45,488 x that), and
// 32-bit accumulation gives the same integers.
constexpr int32_t kIdct32Max = 8190;
// One product of the transform.  In the 32-bit form every operand fits in 24
// signed bits (matrix entries below 2^13, inputs at most kIdct32Max, first-pass
// outputs at most 45,488 x 8190 / 2^13 + 1) and every product in 32, so the
// full-rate 24-bit multiply gives the exact product (a 32-bit multiply issues
// at a quarter of the rate).
template <typename Acc>
__device__ __forceinline__ Acc idct_mul(int32_t c, int32_t x) {
  if constexpr (sizeof(Acc) == 4) return __mul24(c, x);
  else return static_cast<Acc>(c) * x;
}
template <int N, typename Acc>
__device__ __forceinline__ void idct_blocks(WaveScratch& ws, const LdsI32* C, int base, int nblk, uint32_t nzm,
                                            uint32_t acm, int lane) {
  constexpr int NN = N * N;
  const int total = nblk * NN;
  const int32_t c0 = C[0];
  for (int e = lane; e < total; e += 64) {
    const int b = e / NN, local = e - b * NN;
    if ((acm >> b) & 1u) {
      const int i = local / N, l = local % N;
      const int32_t* src = ws.coef + base + b * NN + l;
      Acc acc = 0;
#pragma unroll
      for (int k = 0; k < N; k++) acc += idct_mul<Acc>(C[k * N + i], src[k * N]);
      ws.tmp[base + e] = static_cast<int32_t>((acc + 4096) >> 13);
    } else if (((nzm >> b) & 1u) && local == 0) {
      ws.tmp[base + e] = static_cast<int32_t>((static_cast<Acc>(c0) * ws.coef[base + e] + 4096) >> 13);
    }
  }
  wave_sync();
  for (int e = lane; e < total; e += 64) {
    const int b = e / NN, local = e - b * NN;
    if ((acm >> b) & 1u) {
      const int i = local / N, m = local % N;
      const int32_t* t = ws.tmp + base + b * NN + i * N;
      Acc acc = 0;
#pragma unroll
      for (int l = 0; l < N; l++) acc += idct_mul<Acc>(C[l * N + m], t[l]);
      ws.coef[base + e] = static_cast<int32_t>((acc + 4096) >> 13);
    } else if ((nzm >> b) & 1u) {
      ws.coef[base + e] = static_cast<int32_t>((static_cast<Acc>(c0) * ws.tmp[base + b * NN] + 4096) >> 13);
    }
  }
  wave_sync();
}

// cl: the n x n matrix, cc: the 8 x 8 one.
template <typename Acc>
__device__ __forceinline__ void idct_mb(WaveScratch& ws, const LdsI32* cl, const LdsI32* cc, int n, uint32_t nzm,
                                        uint32_t acm, int lane) {
  if (nzm & 0xFFFFu) {
    if (n == 4) idct_blocks<4, Acc>(ws, cl, 0, 16, nzm, acm, lane);
    else if (n == 8) idct_blocks<8, Acc>(ws, cl, 0, 4, nzm, acm, lane);
    else idct_blocks<16, Acc>(ws, cl, 0, 1, nzm, acm, lane);
  }
  if (nzm & 0x30000u) idct_blocks<8, Acc>(ws, cc, 256, 2, nzm >> 16, acm >> 16, lane);
}

// Largest |coefficient| of the macroblock's dequantised residual over the
// wavefront.
__device__ __forceinline__ int32_t wave_max_abs(const WaveScratch& ws, int lane) {
  int32_t m = 0;
  for (int e = lane; e < 384; e += 64) {
    const int32_t v = ws.coef[e];
    m = max(m, v < 0 ? -v : v);
  }
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
  return m;
}

// Dequantisation and inverse transform of a macroblock's residual into
// ws.coef (dense, zero where nothing is coded).  The wavefront stages only
// the matrices it needs in its own LDS (mats: the n x n matrix at 0, the
// 8 x 8 one at 256): no workgroup barrier.
__device__ __forceinline__ void residual_mb_wave(WaveScratch& ws, LdsI32* mats, const uint32_t* __restrict__ coefs,
                                                 const MBRec& r, int n, int qdc, int qac, int lane) {
  const uint64_t masks = load_coefs(ws, coefs, r, n, qdc, qac, lane);
  const uint32_t nzm = static_cast<uint32_t>(masks), acm = static_cast<uint32_t>(masks >> 32);
  if (!nzm) return;
  if (nzm & 0xFFFFu) {
    const int32_t* src = &c_dct[n == 4 ? 0 : (n == 8 ? 1 : 2)][0];
    for (int i = lane; i < n * n; i += 64) mats[i] = src[i];
  }
  if (nzm & 0x30000u) mats[256 + lane] = c_dct[1][lane];
  wave_sync();
  if (wave_max_abs(ws, lane) <= kIdct32Max) idct_mb<int32_t>(ws, mats, mats + 256, n, nzm, acm, lane);
  else idct_mb<int64_t>(ws, mats, mats + 256, n, nzm, acm, lane);
}

__device__ __forceinline__ int clampc(int v, int hi) { return v < 0 ? 0 : (v >= hi ? hi - 1 : v); }

// ============================================================================
// Inter reconstruction: one wavefront per macroblock
// ============================================================================
// Per wavefront: the residual and the inverse transform's work area, and the
// DCT matrices it needs.
struct InterScratch {
  WaveScratch ws;
  int32_t mats[320];
};

// Partition rectangle pi of a partition mode (partRects, model.go:91-96) as
// x | y << 8 | w << 16 | h << 24.
__device__ __forceinline__ uint32_t part_rect(int part, int pi) {
  switch (part) {
    case kPart16x8: return pi ? 0x08100800u : 0x08100000u;
    case kPart8x16: return pi ? 0x10080008u : 0x10080000u;
    case kPart8x8: return 0x08080000u | ((pi & 1) ? 8u : 0u) | ((pi & 2) ? 0x800u : 0u);
    default: return 0x10100000u;
  }
}

// 4-tap luma filter rows (motioncomp.go:30-35) as packed int8 for v_dot4.
__constant__ uint32_t c_taps8[4] = {0x00004000u, 0xFD1237FAu, 0xF82828F8u, 0xFA3712FDu};

// Four samples clamped to [0, 255] and packed into a dword, sample 0 in the
// low byte.  The empty asm statements keep the compiler from fusing a
// preceding shift, the clamp and the packing into v_ashr_pk_u8_i32: in
// this position it produced wrong upper bytes (samples 2 and 3 of a dword,
// seen as chroma mismatches against the oracle), so the fusion is blocked
// outright.
__device__ __forceinline__ uint32_t clamp_pack4(const int32_t (&v)[4]) {
  int32_t a = v[0], b = v[1], c = v[2], d = v[3];
  asm("" : "+v"(a), "+v"(b), "+v"(c), "+v"(d));
  return static_cast<uint32_t>(clamp255(a)) | (static_cast<uint32_t>(clamp255(b)) << 8) |
         (static_cast<uint32_t>(clamp255(c)) << 16) | (static_cast<uint32_t>(clamp255(d)) << 24);
}

// Two samples clamped to [0, 255] and packed into 16 bits, sample 0 in the
// low byte (the fusion blocked as in clamp_pack4).
__device__ __forceinline__ uint16_t clamp_pack2(int32_t a, int32_t b) {
  asm("" : "+v"(a), "+v"(b));
  return static_cast<uint16_t>(static_cast<uint32_t>(clamp255(a)) | (static_cast<uint32_t>(clamp255(b)) << 8));
}

// Bytes col .. col + 4K - 1 of a plane row of width w as K dwords, each
// column clamped to [0, w) as clampCoord clamps it (motioncomp.go:37-45).
// col is a multiple of 4, so a span inside the row is K aligned dword
// loads; a span over an edge (lanes at the plane's left or right border)
// is gathered byte by byte.
// (Returned by value: output parameters of two inlined calls that the
// compiler merges would become a pointer into scratch.)
template <int K>
struct Span {
  uint32_t d[K];
};
template <int K>
__device__ __forceinline__ Span<K> load_span(const uint8_t* __restrict__ row, int col, int w) {
  Span<K> out;
  uint32_t(&d)[K] = out.d;
  if (col >= 0 && col + 4 * K <= w) {
    const uint32_t* p = reinterpret_cast<const uint32_t*>(row + col);
#pragma unroll
    for (int k = 0; k < K; k++) d[k] = p[k];
  } else {
#pragma unroll
    for (int k = 0; k < K; k++) {
      uint32_t v = 0;
#pragma unroll
      for (int b = 0; b < 4; b++) v |= static_cast<uint32_t>(row[clampc(col + 4 * k + b, w)]) << (8 * b);
      d[k] = v;
    }
  }
  return out;
}

// Signed 4 x int8 dot product of a with taps, accumulator 0.  Asked for with
// saturation, it is v_dot4_i32_i8's three-operand form with the zero as an
// inline constant; without, it is v_dot4c_i32_i8, whose accumulator is its
// destination, and every product costs a v_mov of the zero first.  The sums
// here are far inside int32 (at most 4 x 128 x 128), so saturation never acts.
__device__ __forceinline__ int32_t sdot4z(uint32_t a, uint32_t taps) {
  return __builtin_amdgcn_sdot4(static_cast<int32_t>(a), static_cast<int32_t>(taps), 0, true);
}

// Horizontal 4-tap sums of samples 0-3 from a row segment d0..d2 whose
// first tap is byte sh, on samples biased by -128 (each sum is 128 * 64 low).
__device__ __forceinline__ void luma_hsum4(const Span<3>& s, int sh, int32_t tx, int32_t (&h)[4]) {
  const uint32_t e0 = __builtin_amdgcn_alignbyte(s.d[1], s.d[0], sh) ^ 0x80808080u;
  const uint32_t e1 = __builtin_amdgcn_alignbyte(s.d[2], s.d[1], sh) ^ 0x80808080u;
  const uint32_t t = static_cast<uint32_t>(tx);
  h[0] = sdot4z(e0, t);
  h[1] = sdot4z(__builtin_amdgcn_alignbyte(e1, e0, 1), t);
  h[2] = sdot4z(__builtin_amdgcn_alignbyte(e1, e0, 2), t);
  h[3] = sdot4z(__builtin_amdgcn_alignbyte(e1, e0, 3), t);
}

// Motion compensation of one partition, its residual added and the blocks
// of the write mask stored (decode_inter.go:64-152, motioncomp.go:37-97).
// Luma: lane = 4 samples of partition row lane / 4.  Chroma: lanes 0-31 Cb,
// 32-63 Cr, 2 samples of one row each.  Every reference load of the
// partition is issued before any sample is computed.
//  - Luma: the row segment of a lane's samples is read as aligned dwords
//    and shifted once by the partition's byte phase (the same for every
//    lane, as output columns step by 4); sample j's taps are then bytes
//    j..j+3.  Samples are biased by -128 into signed bytes, each horizontal
//    tap sum is one v_dot4, and the bias returns as 128 * 64 per tap row
//    (the taps sum to 64).  Rows are clamped to the padded plane.  The
//    vertical sum and (v + 2048) >> 12 follow; a {0, 64, 0, 0} tap row
//    makes that equal the reference's 1-D forms, (v + 32) >> 6, and the
//    full-pel case takes the bytes directly.
//  - Chroma: the eighth-pel bilinear sum (8 - fy)((8 - fx) a0 + fx a1) +
//    fy((8 - fx) b0 + fx b1) as two chained v_dot4, one per row, with the
//    row weight folded into the byte taps (each product of weights is at
//    most 64); the second sample's taps are the first's moved up one byte,
//    so both read the same shifted dwords.  (sum + 32) >> 6 is the
//    reference's rounding.
// The reference rows of one lane (luma: up to four 12-byte spans; chroma:
// one 8-byte span of each of two rows).
struct McRows {
  Span<3> q0, q1, q2, q3;
  Span<2> ca, cb;
};

__device__ __forceinline__ void mc_load(const Planes& src, int mbx, int mby, int px, int py, int mvx, int mvy,
                                        int lane, McRows& m) {
  // ---- luma loads
  const int fx = mvx & 3, fy = mvy & 3;
  const int lr = lane >> 2, lc = (lane & 3) * 4;
  const int y = py + lr, x = px + lc;  // macroblock-relative
  const int sy = mby * 16 + y + (mvy >> 2);
  const int sx0 = mbx * 16 + px + (mvx >> 2);
  const bool copy = (fx | fy) == 0;
  const int sft = copy ? sx0 : sx0 - 1;  // first byte read of the partition row
  const int sh = sft & 3;
  const int col = (sft & ~3) + lc;
  // (row offsets by 24-bit multiplies: padded planes are far below 2^24
  // rows and columns)
  auto lrow = [&](int yy) { return src.y + __umul24(static_cast<uint32_t>(clampc(yy, src.ph)), src.pw); };
  if (fy == 0) {  // one row (the full-pel copy uses its first 8 bytes)
    m.q1 = load_span<3>(lrow(sy), col, src.pw);
    asm volatile("");  // keeps the compiler from merging this load with the 2-D path's last one
  } else {
    m.q0 = load_span<3>(lrow(sy - 1), col, src.pw);
    m.q1 = load_span<3>(lrow(sy), col, src.pw);
    m.q2 = load_span<3>(lrow(sy + 1), col, src.pw);
    m.q3 = load_span<3>(lrow(sy + 2), col, src.pw);
  }
  // ---- chroma loads: the dword-aligned 8 bytes holding the lane's first
  // sample column and the two after it, in rows csy and csy + 1
  const int cpx = px >> 1, cpy = py >> 1;
  const int pl = lane >> 5, cr = (lane >> 2) & 7, cc = (lane & 3) * 2;
  const int csy = mby * 8 + cpy + cr + (mvy >> 3);
  const int ccol = (mbx * 8 + cpx + cc + (mvx >> 3)) & ~3;
  const uint8_t* plane = pl ? src.cr : src.cb;
  m.ca = load_span<2>(plane + __umul24(static_cast<uint32_t>(clampc(csy, src.ch)), src.cw), ccol, src.cw);
  m.cb = load_span<2>(plane + __umul24(static_cast<uint32_t>(clampc(csy + 1, src.ch)), src.cw), ccol, src.cw);
}

__device__ __forceinline__ void mc_finish(const Planes& rec, const WaveScratch& ws, const MBRec& r, bool has_res,
                                          int lgn, int mbx, int mby, int px, int py, int pwd, int pht, int mvx,
                                          int mvy, int lane, const McRows& m) {
  const int fx = mvx & 3, fy = mvy & 3;
  const int lr = lane >> 2, lc = (lane & 3) * 4;
  const int y = py + lr, x = px + lc;  // macroblock-relative
  const int sx0 = mbx * 16 + px + (mvx >> 2);
  const bool copy = (fx | fy) == 0;
  const int sh = (copy ? sx0 : sx0 - 1) & 3;
  const int cpx = px >> 1, cpy = py >> 1, cpw = pwd >> 1, cph = pht >> 1;
  const int pl = lane >> 5, cr = (lane >> 2) & 7, cc = (lane & 3) * 2;
  const int cy = cpy + cr, cx = cpx + cc;
  const int csh = (mbx * 8 + cpx + cc + (mvx >> 3)) & 3;
  // ---- luma samples
  // (a full-pel partition without a luma residual stores the shifted
  // reference dword as it is)
  const bool lres = has_res && (r.cbp & kCBPLuma);
  int32_t v[4];
  uint32_t direct = 0;
  if (copy) {
    const uint32_t w = __builtin_amdgcn_alignbyte(m.q1.d[1], m.q1.d[0], sh);
    direct = w;
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = static_cast<int32_t>((w >> (8 * j)) & 0xFFu);
  } else {
    const int32_t tx = static_cast<int32_t>(c_taps8[fx]);
    if (fy == 0) {
      int32_t h[4];
      luma_hsum4(m.q1, sh, tx, h);
#pragma unroll
      for (int j = 0; j < 4; j++) v[j] = clamp255((h[j] + (8192 + 32)) >> 6);
    } else {
      const int32_t t0 = c_taps[fy][0], t1 = c_taps[fy][1], t2 = c_taps[fy][2], t3 = c_taps[fy][3];
      int32_t h0[4], h1[4], h2[4], h3[4];
      luma_hsum4(m.q0, sh, tx, h0);
      luma_hsum4(m.q1, sh, tx, h1);
      luma_hsum4(m.q2, sh, tx, h2);
      luma_hsum4(m.q3, sh, tx, h3);
      // |h| < 2^14 and |tap| <= 64: 24-bit products are exact
#pragma unroll
      for (int j = 0; j < 4; j++)
        v[j] = clamp255((__mul24(t0, h0[j]) + __mul24(t1, h1[j]) + __mul24(t2, h2[j]) + __mul24(t3, h3[j]) +
                         (8192 * 64 + 2048)) >> 12);
    }
  }
  const int blk = ((y >> lgn) << (4 - lgn)) + (x >> lgn);
  if (lr < pht && lc < pwd && ((r.wmask >> blk) & 1u)) {
    if (lres) {
      const int m = (1 << lgn) - 1;
      const int32_t* res = ws.coef + (blk << (2 * lgn)) + ((y & m) << lgn) + (x & m);
#pragma unroll
      for (int j = 0; j < 4; j++) v[j] += res[j];
    }
    uint32_t* dst = reinterpret_cast<uint32_t*>(rec.y + (__umul24(static_cast<uint32_t>(mby * 16 + y), rec.pw) +
                                                         static_cast<uint32_t>(mbx * 16 + x)));
    if (sgpr(static_cast<uint32_t>(copy && !lres)))
      *dst = direct;
    else
      *dst = clamp_pack4(v);
  }
  // ---- chroma samples
  if (cr < cph && cc < cpw && ((r.wmask >> (16 + pl)) & 1u)) {
    const int cfx = mvx & 7, cfy = mvy & 7;
    const uint32_t ea = __builtin_amdgcn_alignbyte(m.ca.d[1], m.ca.d[0], csh) ^ 0x80808080u;
    const uint32_t eb = __builtin_amdgcn_alignbyte(m.cb.d[1], m.cb.d[0], csh) ^ 0x80808080u;
    const int32_t ta = (8 - cfy) * (8 - cfx) | ((8 - cfy) * cfx) << 8, tb = cfy * (8 - cfx) | (cfy * cfx) << 8;
    int32_t u0 = __builtin_amdgcn_sdot4(static_cast<int32_t>(eb), tb,
                                        __builtin_amdgcn_sdot4(static_cast<int32_t>(ea), ta, 8192 + 32, false), false) >> 6;
    int32_t u1 = __builtin_amdgcn_sdot4(static_cast<int32_t>(eb), tb << 8,
                                        __builtin_amdgcn_sdot4(static_cast<int32_t>(ea), ta << 8, 8192 + 32, false),
                                        false) >> 6;
    if (has_res && (r.cbp & (pl ? kCBPCr : kCBPCb))) {
      const int32_t* res = ws.coef + 256 + pl * 64 + cy * 8 + cx;
      u0 += res[0];
      u1 += res[1];
    }
    uint8_t* dp = pl ? rec.cr : rec.cb;
    *reinterpret_cast<uint16_t*>(dp + (__umul24(static_cast<uint32_t>(mby * 8 + cy), rec.cw) +
                                       static_cast<uint32_t>(mbx * 8 + cx))) = clamp_pack2(u0, u1);
  }
}

__device__ __forceinline__ void mc_part(const Planes& src, const Planes& rec, const WaveScratch& ws, const MBRec& r,
                                        bool has_res, int lgn, int mbx, int mby, int px, int py, int pwd, int pht,
                                        int mvx, int mvy, int lane) {
  McRows m{};
  mc_load(src, mbx, mby, px, py, mvx, mvy, lane, m);
  mc_finish(rec, ws, r, has_res, lgn, mbx, mby, px, py, pwd, pht, mvx, mvy, lane, m);
}

// One skip or inter macroblock (reconMCMB / reconInterMB / reconInterChroma,
// decode_inter.go:64-152): the residual (dequantised, inverse transformed)
// first, then motion compensation per partition with the residual added;
// only the blocks of the write mask are stored.
__device__ __forceinline__ void recon_inter_mb(const uint8_t* __restrict__ desc, const MBRec r,
                                               const uint32_t* __restrict__ coefs, ConstJob& job, const int mb,
                                               const int mbx, const int mby, InterScratch& S, const int lane) {
  // Everything per macroblock is wave-uniform: scalar registers and loads.
  // Macroblocks a failed tile stream never reached carry kind kKindUnreached.
  const int mbc = job.mbc;
  if (r.kind == 2) {  // intra: only its residual, for the intra launch
    if (r.coef_count > 0) {
      const FrameHdr* hdr = reinterpret_cast<const FrameHdr*>(desc);
      residual_mb_wave(S.ws, (LdsI32*)S.mats, coefs, r, 4 << (r.tx > 2 ? 2 : r.tx), hdr->qdc, hdr->qac, lane);
      wave_sync();
      uint32_t* dst = reinterpret_cast<uint32_t*>(job.res + static_cast<size_t>(mb) * kResPerMB);
#pragma unroll
      for (int k = 0; k < kResPerMB / 128; k++) {
        const int i = 2 * (lane + 64 * k);
        const int32_t a = max(-32768, min(32767, S.ws.coef[i])), b = max(-32768, min(32767, S.ws.coef[i + 1]));
        dst[lane + 64 * k] = (static_cast<uint32_t>(a) & 0xFFFFu) | (static_cast<uint32_t>(b) << 16);
      }
    }
    return;
  }
  if (r.kind > 2) return;
  const bool has_res = r.kind == 1 && r.coef_count > 0;
  // (the reference is chosen by address in constant memory: selecting
  // between two copies of the plane table would put them in scratch)
  const bool golden = sgpr(static_cast<uint32_t>(r.kind == 1 && r.refsel == kRefGolden)) != 0;
  const Planes rec = planes_of(job.rec);
  const Planes src = planes_of(golden ? job.gold : job.ref);
  const int lgn = 2 + (r.tx > 2 ? 2 : r.tx);  // log2 of the luma transform size
  if (has_res) {
    const FrameHdr* hdr = reinterpret_cast<const FrameHdr*>(desc);
    residual_mb_wave(S.ws, (LdsI32*)S.mats, coefs, r, 1 << lgn, hdr->qdc, hdr->qac, lane);
    wave_sync();  // the residual is read across lanes below
  }
  const int part = r.kind == 1 ? r.part : 0;
  const int npart = static_cast<int>((kPartCountNibbles >> (4 * part)) & 15u);
  // (the vectors as named words: indexing the record copy by the partition
  // number would put it in scratch)
  auto mv_word = [&](int k) {
    return static_cast<uint32_t>(static_cast<uint16_t>(r.mv[k][0])) |
           (static_cast<uint32_t>(static_cast<uint16_t>(r.mv[k][1])) << 16);
  };
  const uint32_t mvw0 = mv_word(0), mvw1 = mv_word(1), mvw2 = mv_word(2), mvw3 = mv_word(3);
  for (int pi = 0; pi < npart; pi++) {
    const uint32_t rc = part_rect(part, pi);
    const int px = rc & 0xFF, py = (rc >> 8) & 0xFF, pwd = (rc >> 16) & 0xFF, pht = rc >> 24;
    const uint32_t mvw = sgpr(pi == 0 ? mvw0 : (pi == 1 ? mvw1 : (pi == 2 ? mvw2 : mvw3)));
    const int mvx = static_cast<int16_t>(mvw & 0xFFFFu), mvy = static_cast<int16_t>(mvw >> 16);
    mc_part(src, rec, S.ws, r, has_res, lgn, mbx, mby, px, py, pwd, pht, mvx, mvy, lane);
  }
}

// items[0, njobs): {job, first macroblock group of the job in this launch},
// ascending; the launch has one workgroup per group of kInterGroupMbs
// macroblocks.
__global__ __launch_bounds__(256, 8) void ebiv_recon_inter(const FrameJob* __restrict__ jobs,
                                                           const WorkItem* __restrict__ items, int njobs) {
  __shared__ InterScratch scratch[4];
  const uint32_t g = xcd_contiguous(blockIdx.x, gridDim.x);
  const int lane = lane_id(), wave = static_cast<int>(sgpr(threadIdx.x >> 6));
  // this group's job: the last item whose first group is <= g (one vector
  // load and a ballot per 64 jobs)
  uint32_t k = 0;
  for (int base = 0; base < njobs; base += 64) {
    const int e = base + lane;
    const uint32_t first = e < njobs ? items[e].idx : 0xFFFFFFFFu;
    const uint32_t c = static_cast<uint32_t>(__popcll(__ballot(first <= g)));
    k += c;
    if (c < 64) break;
  }
  const ConstItem& it = ((ConstItem*)items)[sgpr(k - 1)];
  ConstJob& job = job_of(jobs, it.job);
  const int first_mb = static_cast<int>((g - it.idx) * kInterGroupMbs);
  // The wavefront's two macroblocks first_mb + wave and first_mb + wave + 4:
  // both records are loaded before the first is reconstructed, so only the
  // first waits for the work item, job and record loads in a row.
  static_assert(kInterMbsPerWave == 2, "two macroblocks per wavefront");
  const MBRec* __restrict__ recs = job.recs;
  const int nmb = job.mbc * job.mbr;
  const int mb0 = first_mb + wave, mb1 = mb0 + 4;
  if (mb0 >= nmb) return;
  const MBRec r0 = recs[mb0];
  const MBRec r1 = recs[mb1 < nmb ? mb1 : mb0];
  // their grid positions: one division per wavefront (mb1 is 4 columns on)
  const int mbc = job.mbc;
  const int y0 = static_cast<int>(sgpr(static_cast<uint32_t>(mb0 / mbc))), x0 = mb0 - y0 * mbc;
  int x1 = x0 + 4, y1 = y0;
  while (x1 >= mbc) {  // (at most 4 rounds, when the frame is narrower than 4 macroblocks)
    x1 -= mbc;
    y1++;
  }
  const uint32_t pos1 = sgpr(static_cast<uint32_t>(x1 | y1 << 16));  // one register across the first
  recon_inter_mb(job.desc, r0, job.coefs, job, mb0, x0, y0, scratch[wave], lane);
  if (mb1 < nmb)
    recon_inter_mb(job.desc, r1, job.coefs, job, mb1, static_cast<int>(pos1 & 0xFFFF), static_cast<int>(pos1 >> 16),
                   scratch[wave], lane);
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
  int32_t ref[3][68];    // HEVC reference array per plane in flight
  int32_t ext[3][52];
  int32_t dc[3];
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
  // (24-bit multiplies: every operand is a sample, a weight up to 32 or a
  // position below 17 times an angle of at most 32 in magnitude)
  int pos = __mul24(a + 1, c_ang[mode][0]), idx = pos >> 5, fact = pos & 31;
  int32_t s0 = ext[n + b + idx + 1], s1 = ext[n + b + idx + 2];
  return (__mul24(32 - fact, s0) + __mul24(fact, s1) + 16) >> 5;
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

// Edge samples of the macroblocks reconstructed so far in a tile whose
// macroblocks are all intra (a key frame): per MB column the bottom row of
// its latest macroblock, per MB row the right column of its latest
// macroblock, and per column the above-left corner of the next macroblock
// in rows of either parity.  On the anti-diagonal schedule a macroblock's
// neighbours (left, above-left, above, above-right) are exactly these
// entries when it runs, and no macroblock of the same step writes an entry
// another one of that step reads.
// (kFrontMax: the tile width and height in MBs the LDS frontier covers)
struct Frontier {
  uint8_t bot[kFrontMax][32];        // Y[16], Cb[8], Cr[8] of the bottom row
  uint8_t right[kFrontMax][32];      // Y[16], Cb[8], Cr[8] of the right column
  uint8_t corner[2][kFrontMax + 1][4];  // [row & 1][column]: Y, Cb, Cr
};

// Device-coherent word access (relaxed, agent scope: the L2 of another XCD
// holds no stale copy of what these read or write).
__device__ __forceinline__ uint32_t coherent_load(const uint8_t* p) {
  return __hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void coherent_store(uint8_t* p, uint32_t v) {
  __hip_atomic_store(reinterpret_cast<uint32_t*>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
constexpr int kEdgeIn = 1;   // intra_mb: the row above from the frame (a band's first row)
constexpr int kEdgeOut = 2;  // intra_mb: the bottom rows also device-coherent (a band's last row)

// A macroblock record gathered from lanes 0-11 of v (one word each), for
// records loaded ahead by vector loads (which LDS waits do not hold up).
__device__ __forceinline__ MBRec rec_of_lanes(uint32_t v) {
  MBRec r;
  const uint32_t a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 1);
  r.kind = static_cast<uint8_t>(a);
  r.part = static_cast<uint8_t>(a >> 8);
  r.cbp = static_cast<uint8_t>(a >> 16);
  r.tx = static_cast<uint8_t>(a >> 24);
  r.refsel = static_cast<uint8_t>(b);
  r.cmode = static_cast<uint8_t>(b >> 8);
  r.pad0 = static_cast<uint8_t>(b >> 16);
  r.pad1 = static_cast<uint8_t>(b >> 24);
  r.wmask = __builtin_amdgcn_readlane(v, 2);
  r.coef_off = __builtin_amdgcn_readlane(v, 3);
  r.coef_count = __builtin_amdgcn_readlane(v, 4);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t w = __builtin_amdgcn_readlane(v, 5 + k);
    r.mv[k][0] = static_cast<int16_t>(w & 0xFFFFu);
    r.mv[k][1] = static_cast<int16_t>(w >> 16);
  }
  r.lmode[0] = __builtin_amdgcn_readlane(v, 9);
  r.lmode[1] = __builtin_amdgcn_readlane(v, 10);
  r.pad2 = __builtin_amdgcn_readlane(v, 11);
  return r;
}

// w0-w2: this lane's words of the macroblock's residual as the inter launch
// formed it (words lane, lane + 64, lane + 128; zero without coefficients).
// edge (key-frame bands, with fr): kEdgeIn, kEdgeOut.
__device__ __forceinline__ void intra_mb_w(IntraScratch& S, const uint32_t w0, const uint32_t w1, const uint32_t w2,
                                           const MBRec& r, const TileDesc& td, Planes rec, int mbx, int mby,
                                           Frontier* fr, int lane, int edge = 0) {
  const int n = 4 << (r.tx > 2 ? 2 : r.tx);
  WaveScratch& ws = S.ws;
  {
    const uint32_t w[3] = {w0, w1, w2};
#pragma unroll
    for (int k = 0; k < kResPerMB / 128; k++) {
      const int i = 2 * (lane + 64 * k);
      ws.coef[i] = static_cast<int16_t>(w[k] & 0xFFFFu);
      ws.coef[i + 1] = static_cast<int16_t>(w[k] >> 16);
    }
  }
  // Load the context: luma row above (x from -1 to 31) and left column,
  // chroma likewise; from the frontier in a key frame, else from the frame
  // (its inter macroblocks are final, its intra ones older in this launch).
  const int X = mbx * 16, Y = mby * 16;
  if (fr && (edge & kEdgeIn)) {
    // The row above as the band above published it: whole words around the
    // bytes needed (luma x from X - 4, chroma from CX - 4; Y >= 16, so every
    // word lies inside the plane); the left column from the frontier.
    const int rx = mbx - td.x0, ry = mby - td.y0, tw = td.x1 - td.x0;
    const int CX = mbx * 8, CY = mby * 8;
    if (lane < 9) {
      const uint32_t w = coherent_load(rec.y + static_cast<size_t>(Y - 1) * rec.pw + X - 4 + 4 * lane);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int i = 4 * lane - 3 + j;  // S.L[0][i] is x = X - 1 + i
        if (i >= 0) S.L[0][i] = (i >= 17 && rx + 1 >= tw) ? 0 : static_cast<uint8_t>(w >> (8 * j));
      }
    } else if (lane >= 33 && lane < 49) {
      S.L[lane - 32][0] = fr->right[ry][lane - 33];
    }
    const int pl = lane >> 5, l = lane & 31;
    if (l < 3) {
      const uint8_t* cp = pl ? rec.cr : rec.cb;
      const uint32_t w = coherent_load(cp + static_cast<size_t>(CY - 1) * rec.cw + CX - 4 + 4 * l);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int i = 4 * l - 3 + j;
        if (i >= 0) S.C[pl][0][i] = static_cast<uint8_t>(w >> (8 * j));
      }
    } else if (l >= 9 && l < 17) {
      S.C[pl][l - 8][0] = fr->right[ry][16 + 8 * pl + l - 9];
    }
  } else if (fr) {
    const int rx = mbx - td.x0, ry = mby - td.y0, tw = td.x1 - td.x0;
    if (lane < 33) {
      uint8_t v = 0;
      if (lane == 0) v = fr->corner[ry & 1][rx][0];
      else if (lane <= 16) v = fr->bot[rx][lane - 1];
      else if (rx + 1 < tw) v = fr->bot[rx + 1][lane - 17];
      S.L[0][lane] = v;
    } else if (lane < 49) {
      S.L[lane - 32][0] = fr->right[ry][lane - 33];
    }
    const int pl = lane >> 5, l = lane & 31;
    if (l < 9) S.C[pl][0][l] = l == 0 ? fr->corner[ry & 1][rx][1 + pl] : fr->bot[rx][16 + 8 * pl + l - 1];
    else if (l < 17) S.C[pl][l - 8][0] = fr->right[ry][16 + 8 * pl + l - 9];
  } else {
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
        S.C[pl][y + 1][x0 + k + 1] = static_cast<uint8_t>(v);
      }
      uint8_t* dp = pl ? rec.cr : rec.cb;
      *reinterpret_cast<uint16_t*>(dp + static_cast<size_t>(mby * 8 + y) * rec.cw + mbx * 8 + x0) = static_cast<uint16_t>(half);
    }
  }
  wave_sync();
  if (edge & kEdgeOut) {
    // the bottom rows once more, device-coherent, for the band below (the
    // same words as above, under the same write mask)
    if (lane < 4) {
      const int blk = (15 / n) * per + (4 * lane) / n;
      if ((r.wmask >> blk) & 1u) {
        uint32_t w = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) w |= static_cast<uint32_t>(S.L[16][4 * lane + k + 1]) << (8 * k);
        coherent_store(rec.y + static_cast<size_t>(Y + 15) * rec.pw + X + 4 * lane, w);
      }
    } else if (lane < 8) {
      const int pl = (lane - 4) >> 1, k2 = lane & 1;
      if ((r.wmask >> (16 + pl)) & 1u) {
        uint32_t w = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) w |= static_cast<uint32_t>(S.C[pl][8][4 * k2 + k + 1]) << (8 * k);
        coherent_store((pl ? rec.cr : rec.cb) + static_cast<size_t>(mby * 8 + 7) * rec.cw + mbx * 8 + 4 * k2, w);
      }
    }
  }
  if (fr) {  // this macroblock's edges for its right and lower neighbours
    const int rx = mbx - td.x0, ry = mby - td.y0;
    if (lane < 16) {
      fr->bot[rx][lane] = S.L[16][lane + 1];
      fr->right[ry][lane] = S.L[lane + 1][16];
    } else if (lane < 32) {
      const int pl = (lane - 16) >> 3, k = (lane - 16) & 7;
      fr->bot[rx][16 + 8 * pl + k] = S.C[pl][8][k + 1];
      fr->right[ry][16 + 8 * pl + k] = S.C[pl][k + 1][8];
    } else if (lane < 35) {
      const int c = lane - 32;
      fr->corner[(ry + 1) & 1][rx + 1][c] = c == 0 ? S.L[16][16] : S.C[c - 1][8][8];
    }
    wave_sync();
  }
}

// A whole 16x16 intra macroblock (transform size 2, every plane written) on
// the LDS frontier, in three phases instead of intra_mb_w's block chain:
//   A: the reference arrays of luma, Cb and Cr straight from the frontier;
//   B: the angular lines and DC values of all three;
//   C: every lane predicts its 4 luma and 2 chroma samples, adds the
//      residual words it holds (w0: luma row lane / 8, columns 2 (lane % 8)
//      and + 1; w1: the same 8 rows lower; w2: chroma plane lane / 32, row
//      (lane % 32) / 4, columns 2 (lane % 4) and + 1), stores them and
//      writes the frontier entries it holds.
// Prediction and substitution are intra_mb_w's (build_ref_one,
// prep_angular, intra_pixel), so the samples are the same.
__device__ __forceinline__ void intra_mb16(IntraScratch& S, const uint32_t w0, const uint32_t w1, const uint32_t w2,
                                           const MBRec& r, const TileDesc& td, Planes rec, int mbx, int mby,
                                           Frontier& fr, int lane, int edge) {
  const int rx = mbx - td.x0, ry = mby - td.y0, tw = td.x1 - td.x0;
  const int X = mbx * 16, Y = mby * 16, CX = mbx * 8, CY = mby * 8;
  if (edge & kEdgeIn) {
    // the band above's last row into this row's frontier entries: above,
    // above-right (luma) and the corner (no macroblock of this band writes
    // them before this one reads them, and the ones after it overwrite them)
    if (lane < 9) {
      const uint32_t w = coherent_load(rec.y + static_cast<size_t>(Y - 1) * rec.pw + X - 4 + 4 * lane);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int i = 4 * lane - 3 + j;  // x = X - 1 + i
        const uint8_t v = static_cast<uint8_t>(w >> (8 * j));
        if (i == 0) fr.corner[ry & 1][rx][0] = v;
        else if (i >= 1 && i <= 16) fr.bot[rx][i - 1] = v;
        else if (i >= 17 && rx + 1 < tw) fr.bot[rx + 1][i - 17] = v;
      }
    } else if (lane >= 16 && lane < 22) {
      const int pl = (lane - 16) / 3, k = (lane - 16) % 3;
      const uint8_t* cp = pl ? rec.cr : rec.cb;
      const uint32_t w = coherent_load(cp + static_cast<size_t>(CY - 1) * rec.cw + CX - 4 + 4 * k);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int i = 4 * k - 3 + j;
        const uint8_t v = static_cast<uint8_t>(w >> (8 * j));
        if (i == 0) fr.corner[ry & 1][rx][1 + pl] = v;
        else if (i >= 1) fr.bot[rx][16 + 8 * pl + i - 1] = v;
      }
    }
    wave_sync();
  }
  // ---- A: reference arrays.  Each entry is one frontier byte (or the
  // no-neighbour value): build_ref_one's choice of sample written as an
  // offset into the frontier, so that a lane issues all its reads at once
  // instead of one per branch of the choice.
  const int tl = td.x0 * 16, tt = td.y0 * 16, trr = td.x1 * 16;
  const bool av_above = Y > tt, av_left = X > tl;
  const bool av_ar = !(Y <= tt || X + 32 > trr);  // aboveRightAvailable, predict.go:50-58, at the MB origin
  {
    const uint8_t* fb = reinterpret_cast<const uint8_t*>(&fr);
    constexpr int kBot = offsetof(Frontier, bot), kRight = offsetof(Frontier, right),
                  kCorner = offsetof(Frontier, corner);
    const int bot0 = kBot + rx * 32, right0 = kRight + ry * 32, corner0 = kCorner + ((ry & 1) * (kFrontMax + 1) + rx) * 4;
    // plane p (0 luma, 1 Cb, 2 Cr) of n x n, entry i in [0, 4n]: the byte offset
    auto src = [&](int p, int n, int i) -> int {
      const int po = p == 0 ? 0 : 8 + 8 * p;  // the plane's bytes in a frontier entry
      const int above0 = bot0 + po, left0 = right0 + po;
      if (i < 2 * n) {
        const int rr = 2 * n - 1 - i < n ? 2 * n - 1 - i : n - 1;
        return av_left ? left0 + rr : above0;
      }
      if (i == 2 * n) return av_left ? (av_above ? corner0 + p : left0) : above0;
      int j = i - 2 * n - 1;
      if (j >= n && (p != 0 || !av_ar)) j = n - 1;
      return av_above ? (j < n ? above0 + j : bot0 + 32 + (j - n)) : left0;
    };
    const bool none = !av_left && !av_above;
    const int oy = src(0, 16, lane), ocb = src(1, 8, lane < 33 ? lane : 32), ocr = src(2, 8, lane < 33 ? lane : 32);
    const int oy64 = src(0, 16, 64);
    const int32_t vy = fb[oy], vcb = fb[ocb], vcr = fb[ocr], vy64 = fb[oy64];
    S.ref[0][lane] = none ? kDCPred : vy;
    if (lane == 0) S.ref[0][64] = none ? kDCPred : vy64;
    if (lane < 33) {
      S.ref[1][lane] = none ? kDCPred : vcb;
      S.ref[2][lane] = none ? kDCPred : vcr;
    }
  }
  wave_sync();
  // ---- B: angular lines and DC values
  const int lmode = r.lmode[0] & 15, cmode = r.cmode;
  prep_angular(S.ref[0], S.ext[0], &S.dc[0], 16, lmode, lane);
  {
    const int pl = lane >> 5;
    prep_angular(S.ref[1 + pl], S.ext[1 + pl], &S.dc[1 + pl], 8, cmode, lane & 31);
  }
  wave_sync();
  // ---- C: samples, frame, frontier
  // (sample pairs: rows q and q + 8 of luma, then chroma)
  const int q = lane >> 3, c = 2 * (lane & 7);
  const int pl = lane >> 5, cr = (lane & 31) >> 2, cc = 2 * (lane & 3);
#pragma unroll
  for (int h = 0; h < 3; h++) {
    const bool luma = h < 2;
    const uint32_t w = h == 0 ? w0 : (h == 1 ? w1 : w2);
    const int rr = luma ? q + 8 * h : cr, c0 = luma ? c : cc, n = luma ? 16 : 8, mode = luma ? lmode : cmode;
    const int pi = luma ? 0 : 1 + pl;
    const int32_t a0 = clamp255(intra_pixel(S.ref[pi], S.ext[pi], S.dc[pi], n, mode, rr, c0) +
                                static_cast<int16_t>(w & 0xFFFFu));
    const int32_t a1 = clamp255(intra_pixel(S.ref[pi], S.ext[pi], S.dc[pi], n, mode, rr, c0 + 1) +
                                static_cast<int16_t>(w >> 16));
    uint8_t* dp = luma ? rec.y + static_cast<size_t>(Y + rr) * rec.pw + X
                       : (pl ? rec.cr : rec.cb) + static_cast<size_t>(CY + rr) * rec.cw + CX;
    *reinterpret_cast<uint16_t*>(dp + c0) = static_cast<uint16_t>(a0 | (a1 << 8));
    const int fo = luma ? 0 : 16 + 8 * pl;  // the plane's offset in a frontier entry
    if (rr == n - 1) {  // the bottom row
      fr.bot[rx][fo + c0] = static_cast<uint8_t>(a0);
      fr.bot[rx][fo + c0 + 1] = static_cast<uint8_t>(a1);
    }
    if (c0 + 1 == n - 1) {  // the right column (and the corner)
      fr.right[ry][fo + rr] = static_cast<uint8_t>(a1);
      if (rr == n - 1) fr.corner[(ry + 1) & 1][rx + 1][luma ? 0 : 1 + pl] = static_cast<uint8_t>(a1);
    }
  }
  wave_sync();
  if (edge & kEdgeOut) {
    // the bottom rows once more, device-coherent, for the band below
    if (lane < 8) {
      uint32_t w = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) w |= static_cast<uint32_t>(fr.bot[rx][4 * lane + k]) << (8 * k);
      if (lane < 4) {
        coherent_store(rec.y + static_cast<size_t>(Y + 15) * rec.pw + X + 4 * lane, w);
      } else {
        const int pl = (lane - 4) >> 1, k2 = lane & 1;
        coherent_store((pl ? rec.cr : rec.cb) + static_cast<size_t>(CY + 7) * rec.cw + CX + 4 * k2, w);
      }
    }
  }
}

// res: the macroblock's residual as the inter launch formed it.
__device__ __forceinline__ void intra_mb(IntraScratch& S, const int16_t* __restrict__ res, const MBRec& r,
                                         const TileDesc& td, Planes rec, int mbx, int mby, Frontier* fr, int lane) {
  uint32_t w0 = 0, w1 = 0, w2 = 0;
  if (r.coef_count > 0) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(res);
    w0 = src[lane];
    w1 = src[lane + 64];
    w2 = src[lane + 128];
  }
  intra_mb_w(S, w0, w1, w2, r, td, rec, mbx, mby, fr, lane);
}

// ---------------------------------------------------------------------------
// Intra reconstruction by block levels: key-frame tiles and tiles dense with
// intra macroblocks.
//
// Go reconstructs a tile's macroblocks in raster order and each macroblock's
// blocks in raster order (decode_intra.go:121-187); a block reads the
// reconstructed samples to its left, above, above-right (when
// aboveRightAvailable, predict.go:50-58) and above-left.  Give macroblock
// (x, y) of the tile the rank t = x + 2y and each of its units (a luma
// block, or the chroma pair) the level 6t + local, with local
//   4x4 block (bx, by): bx + 2by   8x8 block: 0, 2, 4, 5 in raster order
//   16x16: 4                       chroma (Cb and Cr, one mode): 0.
// Every sample a unit reads belongs to a unit of a lower level: within a
// macroblock the raster dependencies climb the local order; the left and
// above-right macroblocks have rank t - 1 (levels up to 6t + 3, below every
// local of the right half of a macroblock's first unit row that reads
// them), the above macroblock t - 2 and the above-left t - 3; and no unit
// reads a sample of the next macroblock, whose levels start at 6t + 6.  So
// the units of one level are independent and run in one parallel round,
// and a 30 x 23-macroblock tile needs at most 454 rounds where the
// macroblock wavefront needs 75 steps of up to 16 serial blocks.
//
// Wavefront w of W owns the tile's macroblock rows w, w + W, w + 2W, ...
// and sweeps them: in levels 6t .. 6t + 5 its row y with 2y <= t <= 2y + tw
// works on macroblock x = t - 2y at locals 0-5 and on its left neighbour at
// locals 6-9 (the last 4x4 blocks).  It has at most one such row when
// th <= W or tw < 2W.  Its lanes have fixed roles: 0-31 the chroma pair,
// 32-47 the macroblock's two 4x4 blocks (4 lanes each) or its 8x8 block,
// 48-55 the left neighbour's two 4x4 blocks, all 64 a 16x16 block: one lane
// per 4-sample row segment.  A unit's lanes first read its reference array
// (predict.go:63-90) into the wavefront's LDS, then predict from it and
// write the frame and the frontier.
//
// The context lives in LDS as a frontier: per luma column the bottom row of
// the latest unit reconstructed there, per luma row the right column of the
// latest one, per 4x4 cell its bottom-right sample (the corner of the cells
// below-right of it), and the same per macroblock for chroma.  A unit's
// above row is the bottom row of the unit directly above it (no unit below
// it in those columns is reconstructed yet: each depends on it or on its
// right neighbour), its left column likewise; units of one level write
// frontier entries no other unit of that level reads.  Samples of inter
// macroblocks, final before this launch, are read from the frame.
//
// A macroblock's residual (768 bytes) is loaded during the rank before its
// own into a 3-slot LDS ring of the wavefront (slot: its rank mod 3) and
// parked there before that rank's last frame stores: the compiler waits for
// every outstanding vector memory operation, stores included, when it
// waits for a load while stores are pending, so the wait falls on stores
// at least one level old.
constexpr int kCellMaxMbs = 1024;   // tile macroblocks the cell path takes
constexpr int kCellMaxDim = 64;     // tile width and height in macroblocks
constexpr int kCellStride = 6;      // levels per macroblock rank
constexpr int kCellRankWords = (3 * kCellMaxDim) / 32 + 1;
constexpr int kCellRefPerLane = 5;  // reference bytes per lane of a unit (4x4: 17 over 4 lanes)

struct CellFront {
  uint8_t A[kCellMaxDim * 16];      // luma: bottom row of the latest unit per column
  uint8_t Lc[kCellMaxDim * 16];     // luma: right column of the latest unit per row
  uint8_t BR[kCellMaxMbs * 16];     // luma: bottom-right sample of each 4x4 cell
  uint8_t Ac[2][kCellMaxDim * 8];   // chroma: the same per plane
  uint8_t Lcc[2][kCellMaxDim * 8];
  uint8_t BRc[2][kCellMaxMbs];      // chroma: bottom-right sample of each macroblock
  uint2 lmode[kCellMaxMbs];         // luma modes (MBRec::lmode)
  uint32_t info[kCellMaxMbs];       // wmask | transform index << 18 | residual << 20 | chroma mode << 24
  uint32_t intra[kCellMaxMbs / 32]; // intra macroblocks (tile raster index)
  uint32_t ranks[kCellRankWords];   // ranks holding an intra macroblock
};
struct CellWave {
  uint32_t res[3][kResPerMB / 2];    // residual ring, the macroblock of rank t in slot t % 3
  uint8_t ref[64 * kCellRefPerLane];  // reference arrays of this level's units
};
__host__ __device__ constexpr size_t cell_lds_bytes(int waves) {
  return sizeof(CellFront) + static_cast<size_t>(waves) * sizeof(CellWave);
}
// Tiles the cell path takes with `waves` wavefronts.
__host__ __device__ constexpr bool cell_tile_fits(int tw, int th, int waves) {
  return tw <= kCellMaxDim && th <= kCellMaxDim && tw * th <= kCellMaxMbs && (th <= waves || tw < 2 * waves);
}

// Luma unit `slot` (0 or 1) of local k in a macroblock with transform index
// tx, or -1: returns type << 4 | by << 2 | bx.
__device__ __forceinline__ int cell_luma(int tx, int k, int slot) {
  if (tx == 0) {  // 4x4: bx + 2by = k
    const int by = (k >> 1) - slot, bx = (k & 1) + 2 * slot;
    return (by >= 0 && by <= 3) ? (by << 2 | bx) : -1;
  }
  if (slot == 1) return -1;
  if (tx == 1) {  // 8x8
    switch (k) {
      case 0: return 1 << 4;
      case 2: return 1 << 4 | 1;
      case 4: return 1 << 4 | 1 << 2;
      case 5: return 1 << 4 | 1 << 2 | 1;
      default: return -1;
    }
  }
  return k == 4 ? 2 << 4 : -1;  // 16x16
}
// Synchronisation of the block levels orders LDS only: no sample written to
// the frame in this launch is read back in it (inter samples are final
// before it, intra ones travel through the frontier), so the frame stores
// and the residual loads stay in flight across levels where a fence or
// __syncthreads() would wait for them.
__device__ __forceinline__ void cell_wave_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ void cell_block_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
// kAngModes in arithmetic form (a lookup by a per-lane mode is a memory
// load): D45..D113 repeat angle 32, 13, -13 (the negative one with inverse
// angle -630), the first three vertical.
__host__ __device__ constexpr int ang_angle(int m) {
  return m < kModeD45 ? 0 : ((m - kModeD45) % 3 == 0 ? 32 : ((m - kModeD45) % 3 == 1 ? 13 : -13));
}
__host__ __device__ constexpr int ang_inv(int m) { return m >= kModeD45 && (m - kModeD45) % 3 == 2 ? -630 : 0; }
__host__ __device__ constexpr bool ang_vert(int m) { return m >= kModeD45 && m < kModeD45 + 3; }
constexpr bool ang_form_matches(int m) {
  return m == kNumIntraModes || (ang_angle(m) == kAngModes[m].angle && ang_inv(m) == kAngModes[m].inv_angle &&
                                 ang_vert(m) == kAngModes[m].vertical && ang_form_matches(m + 1));
}
static_assert(ang_form_matches(kModeD45), "angular parameters");
__device__ void intra_cells(uint8_t* __restrict__ lds, const MBRec* __restrict__ recs, const int16_t* __restrict__ res,
                            const TileDesc& td, const int mbc, const Planes rec, const uint32_t mbs_done,
                            const int lane, const int wave, const int nwaves) {
  CellFront& F = *reinterpret_cast<CellFront*>(lds);
  CellWave& CW = reinterpret_cast<CellWave*>(lds + sizeof(CellFront))[wave];
  const int tw = td.x1 - td.x0, th = td.y1 - td.y0;
  const int tpx = tw * 16;  // tile width in luma samples
  const int tid = static_cast<int>(threadIdx.x), nthr = static_cast<int>(blockDim.x);
  for (int i = tid; i < kCellMaxMbs / 32; i += nthr) F.intra[i] = 0;
  for (int i = tid; i < kCellRankWords; i += nthr) F.ranks[i] = 0;
  __syncthreads();
  const int nmb = tw * th;
  for (int i = tid; i < nmb; i += nthr) {
    const int ry = i / tw, rx = i - ry * tw;
    const MBRec& r = recs[(td.y0 + ry) * mbc + td.x0 + rx];
    const uint4 w = *reinterpret_cast<const uint4*>(&r);  // kind..tx, refsel..pad1, wmask, coef_off
    const uint32_t ncoef = reinterpret_cast<const uint32_t*>(&r)[4];
    F.lmode[i] = make_uint2(r.lmode[0], r.lmode[1]);
    F.info[i] = (w.z & 0x3FFFFu) | min((w.x >> 24) & 0xFFu, 2u) << 18 | (ncoef > 0 ? 1u : 0u) << 20 |
                ((w.y >> 8) & 0xFFu) << 24;
    if (static_cast<uint32_t>(i) < mbs_done && (w.x & 0xFFu) == 2) {
      atomicOr(&F.intra[i >> 5], 1u << (i & 31));
      const int t = rx + 2 * ry;
      atomicOr(&F.ranks[t >> 5], 1u << (t & 31));
    }
  }
  __syncthreads();
  auto is_intra = [&](int li) { return (F.intra[li >> 5] >> (li & 31)) & 1u; };
  auto rank_has = [&](int t) { return t >= 0 && ((F.ranks[t >> 5] >> (t & 31)) & 1u); };
  const int gxl = td.x0 * 16, gyl = td.y0 * 16, gxc = td.x0 * 8, gyc = td.y0 * 8;
  // This wavefront's row at rank t (-1: none): the last of its rows (those
  // below th) with 2y <= t, if t - 2y <= tw.
  const int y_last = wave < th ? wave + nwaves * ((th - 1 - wave) / nwaves) : -1;
  auto row_at = [&](int t) -> int {
    if (y_last < 0 || t < 2 * wave) return -1;
    const int y = min(wave + nwaves * ((t / 2 - wave) / nwaves), y_last);
    return t - 2 * y <= tw ? y : -1;
  };
  // Its macroblock of rank t (tile raster index) if that has a residual, else -1.
  auto res_mb = [&](int t) -> int {
    const int y = row_at(t);
    if (y < 0 || t - 2 * y >= tw) return -1;
    const int li = y * tw + t - 2 * y;
    return (is_intra(li) && ((F.info[li] >> 20) & 1u)) ? li : -1;
  };
  auto res_src = [&](int li) {
    const int ry = li / tw, rx = li - ry * tw;
    return reinterpret_cast<const uint3*>(res + static_cast<size_t>((td.y0 + ry) * mbc + td.x0 + rx) * kResPerMB) + lane;
  };
  {
    const int li = res_mb(0);
    if (li >= 0) reinterpret_cast<uint3*>(CW.res[0])[lane] = *res_src(li);
  }
  const int tmax = (tw - 1) + 2 * (th - 1);
  for (int t1 = 0; t1 <= tmax + 1; t1++) {
    // the next rank's residual, loaded now and parked before this rank's last stores
    const int li_next = res_mb(t1 + 1);
    uint32_t* const park = CW.res[(t1 + 1) % 3];
    if (!rank_has(t1) && !rank_has(t1 - 1)) {  // (uniform) no unit in these 6 levels
      if (li_next >= 0) reinterpret_cast<uint3*>(park)[lane] = *res_src(li_next);
      continue;
    }
    const uint3 nx = li_next >= 0 ? *res_src(li_next) : make_uint3(0u, 0u, 0u);
    // this wavefront's macroblocks: x (rank t1) and its left neighbour (t1 - 1)
    const int y = row_at(t1);
    const int x = y >= 0 ? t1 - 2 * y : tw;
    const int lic = y >= 0 ? y * tw + min(x, tw - 1) : 0;
    const int lip = y >= 0 ? y * tw + max(x - 1, 0) : 0;
    const bool has_cur = y >= 0 && x < tw && is_intra(lic);
    const bool has_prev = y >= 0 && x >= 1 && is_intra(lip);
    const uint32_t info_c = F.info[lic], info_p = F.info[lip];
    const uint2 lm_c = F.lmode[lic], lm_p = F.lmode[lip];
    const int txc = static_cast<int>((info_c >> 18) & 3u), txp = static_cast<int>((info_p >> 18) & 3u);
    for (int k1 = 0; k1 < kCellStride; k1++) {
      // ---- this lane's unit
      int u = -1, qi = 0, nq = 4;
      bool prev = false;
      if (has_cur) {
        if (txc == 2 && k1 == 4) {
          u = 2 << 4; qi = lane; nq = 64;
        } else if (lane < 32) {
          if (k1 == 0) { u = 3 << 4; qi = lane; nq = 32; }
        } else if (lane < 48) {
          if (txc == 1) { u = cell_luma(1, k1, 0); qi = lane - 32; nq = 16; }
          else if (txc == 0 && lane < 40) { u = cell_luma(0, k1, (lane - 32) >> 2); qi = (lane - 32) & 3; }
        }
      }
      if (has_prev && txp == 0 && k1 <= 3 && lane >= 48 && lane < 56) {
        u = cell_luma(0, k1 + kCellStride, (lane - 48) >> 2); qi = (lane - 48) & 3; prev = true;
      }
      const bool active = u >= 0;
      const int li = prev ? lip : lic;
      const int ry = li / tw, rx = li - ry * tw;
      const uint32_t info = prev ? info_p : info_c;
      const uint2 lm = prev ? lm_p : lm_c;
      const int type = active ? u >> 4 : 0, bxu = u & 3, byu = (u >> 2) & 3;
      const bool chroma = type == 3;
      const int n = chroma ? 8 : (4 << type);
      const int sh = chroma ? 3 : 4;
      const int X0 = chroma ? rx * 8 : rx * 16 + bxu * n, Y0 = chroma ? ry * 8 : ry * 16 + byu * n;
      const bool av_above = Y0 > 0, av_left = X0 > 0;
      bool av_ar = false;  // aboveRightAvailable, predict.go:50-58 (never for chroma)
      if (!chroma && Y0 > 0 && X0 + 2 * n <= tpx) av_ar = Y0 == ry * 16 || X0 + 2 * n <= rx * 16 + 16;
      // this lane's row segment: plane, row, first column
      const int pl = chroma ? qi >> 4 : 0;
      const int rr = chroma ? (qi & 15) >> 1 : qi / (n >> 2);
      const int c0 = chroma ? (qi & 1) * 4 : (qi - rr * (n >> 2)) * 4;
      const int blk = chroma ? 16 + pl : byu * (16 / n) + bxu;
      // its residual, from the ring
      const uint32_t* rsl = CW.res[(prev ? t1 + 2 : t1) % 3];  // (t1 - 1) % 3 for the neighbour
      const uint2 rv = active && ((info >> 20) & 1u)
                           ? *reinterpret_cast<const uint2*>(rsl + (chroma ? 128 + pl * 32 + rr * 4 + (c0 >> 1)
                                                                          : (blk * n * n + rr * n + c0) >> 1))
                           : make_uint2(0u, 0u);
      // ---- the reference array: entry e of R = 4n + 1 (per chroma plane),
      // with the substitutions of predict.go:134-156 in closed form
      uint8_t* ref = CW.ref + (lane - qi) * kCellRefPerLane;
      const int R = 4 * n + 1, Rt = chroma ? 2 * R : R;
      // (branch-free, so every entry's LDS reads issue together; an entry not
      // needed reads the unit's own first sample position, in bounds)
      const bool any_av = av_left || av_above;
      const uint8_t* fl = reinterpret_cast<const uint8_t*>(&F);
      int32_t got[kCellRefPerLane];
      uint32_t iw[kCellRefPerLane];
      int lsh[kCellRefPerLane];
      bool need[kCellRefPerLane];
      const uint8_t* gp[kCellRefPerLane];
#pragma unroll
      for (int k = 0; k < kCellRefPerLane; k++) {
        const int e = qi + k * nq;
        need[k] = active && e < Rt && any_av;
        const int epl = e >= R ? 1 : 0, i = e - epl * R;
        int dx = 0, dy = 0;
        if (need[k]) {
          if (i < 2 * n) {
            dx = av_left ? -1 : 0;
            dy = av_left ? min(2 * n - 1 - i, n - 1) : -1;
          } else if (i == 2 * n) {
            dx = av_left ? -1 : 0;
            dy = av_left ? (av_above ? -1 : 0) : -1;
          } else if (!av_above) {
            dx = -1;
          } else {
            const int j = i - 2 * n - 1;
            dx = (j >= n && !av_ar) ? n - 1 : j;
            dy = -1;
          }
        }
        const int x2 = X0 + dx, y2 = Y0 + dy;
        const int mli = (y2 >> sh) * tw + (x2 >> sh);
        iw[k] = F.intra[mli >> 5];
        lsh[k] = mli & 31;
        const bool corner = dx < 0 && dy < 0;
        size_t lo;
        if (chroma) {
          lo = corner ? offsetof(CellFront, BRc) + epl * kCellMaxMbs + mli
                      : (dx < 0 ? offsetof(CellFront, Lcc) + epl * (kCellMaxDim * 8) + y2
                                : offsetof(CellFront, Ac) + epl * (kCellMaxDim * 8) + x2);
        } else {
          lo = corner ? offsetof(CellFront, BR) + (y2 >> 2) * (tw * 4) + (x2 >> 2)
                      : (dx < 0 ? offsetof(CellFront, Lc) + y2 : offsetof(CellFront, A) + x2);
        }
        got[k] = fl[lo];
        gp[k] = chroma ? (epl ? rec.cr : rec.cb) + static_cast<size_t>(gyc + y2) * static_cast<size_t>(rec.cw) + (gxc + x2)
                       : rec.y + static_cast<size_t>(gyl + y2) * static_cast<size_t>(rec.pw) + (gxl + x2);
      }
      // Samples of inter macroblocks from the frame: loaded for every entry
      // when any lane of the wavefront needs one (never in a key frame), so
      // the loads issue together.
      bool from_frame[kCellRefPerLane];
      bool any_frame = false;
#pragma unroll
      for (int k = 0; k < kCellRefPerLane; k++) {
        from_frame[k] = need[k] && !((iw[k] >> lsh[k]) & 1u);
        any_frame |= from_frame[k];
      }
      if (__ballot(any_frame) != 0) {
        int32_t fv[kCellRefPerLane];
#pragma unroll
        for (int k = 0; k < kCellRefPerLane; k++) fv[k] = *gp[k];
#pragma unroll
        for (int k = 0; k < kCellRefPerLane; k++) got[k] = from_frame[k] ? fv[k] : got[k];
      }
#pragma unroll
      for (int k = 0; k < kCellRefPerLane; k++) {
        const int e = qi + k * nq;
        if (active && e < Rt) ref[e] = static_cast<uint8_t>(any_av ? got[k] : kDCPred);
      }
      cell_wave_sync();
      const bool writes = active && ((info >> blk) & 1u);  // reached by its tile stream
      uint32_t word = 0;
      if (writes) {
        const uint8_t* rp = ref + pl * R;
        const int mode = min(chroma ? static_cast<int>(info >> 24)
                                    : static_cast<int>(((blk < 8 ? lm.x >> (4 * blk) : lm.y >> (4 * (blk - 8)))) & 15u),
                             kNumIntraModes - 1);  // (a decoded block's mode is always in range)
        // Every mode as (w0 ref[i0] + w1 ref[i1] + w2 ref[i2] + 16) >> 5,
        // clamped (a no-op except for TrueMotion): DC from the mean, V/H
        // w0 = 32, TM 32, 32, -32 (top + left - corner), angular the two
        // taps of predict.go:162-203.
        int32_t dc = 0;
        if (mode == kModeDC) {  // the n left samples nearest the corner and the first n above
#pragma unroll
          for (int i = 0; i < 16; i++) {
            const int32_t two = rp[min(n + i, 2 * n - 1)] + rp[min(2 * n + 1 + i, 3 * n)];
            dc += i < n ? two : 0;
          }
          dc = (dc + n) / (2 * n);
        }
        const int angle = ang_angle(mode), inv = ang_inv(mode);
        const bool vert = ang_vert(mode);
        // main line ext[kk] of the angular modes as a reference index (kk in
        // [0, 2n]), its projected side for kk < 0; -1 past 2n (weight 0)
        auto ext_index = [&](int kk) -> int {
          if (kk > 2 * n) return -1;
          if (kk >= 0) return vert ? 2 * n + kk : 2 * n - kk;
          const int idx = (kk * inv + 128) >> 8;
          return vert ? 2 * n - idx : 2 * n + idx;
        };
#pragma unroll
        for (int j = 0; j < 4; j++) {
          int i0 = 0, i1 = 0, i2 = 0, w0 = 32, w1 = 0, w2 = 0;
          if (mode == kModeV) {
            i0 = 2 * n + 1 + c0 + j;
          } else if (mode == kModeH) {
            i0 = 2 * n - 1 - rr;
          } else if (mode == kModeTM) {
            i0 = 2 * n + 1 + c0 + j;
            i1 = 2 * n - 1 - rr;
            i2 = 2 * n;
            w1 = 32;
            w2 = -32;
          } else if (mode != kModeDC) {
            const int a2 = vert ? rr : c0 + j, b2 = vert ? c0 + j : rr;
            const int pos = (a2 + 1) * angle, idx = pos >> 5, f = pos & 31;
            i0 = ext_index(b2 + idx + 1);
            i1 = ext_index(b2 + idx + 2);
            w0 = i0 < 0 ? 0 : 32 - f;
            w1 = i1 < 0 ? 0 : f;
            i0 = max(i0, 0);
            i1 = max(i1, 0);
          }
          const int32_t acc = mode == kModeDC ? 32 * dc + 16 : w0 * rp[i0] + w1 * rp[i1] + w2 * rp[i2] + 16;
          const uint32_t rw16 = (j < 2 ? rv.x : rv.y) >> (16 * (j & 1));
          const int32_t val = clamp255(clamp255(acc >> 5) + static_cast<int16_t>(rw16 & 0xFFFFu));
          word |= static_cast<uint32_t>(val) << (8 * j);
        }
      }
      if (k1 == kCellStride - 1) reinterpret_cast<uint3*>(park)[lane] = nx;  // before this rank's last stores
      if (writes) {
        const uint8_t last = static_cast<uint8_t>(word >> 24);
        if (chroma) {
          uint8_t* dp = pl ? rec.cr : rec.cb;
          *reinterpret_cast<uint32_t*>(dp + static_cast<size_t>(gyc + Y0 + rr) * static_cast<size_t>(rec.cw) +
                                       static_cast<size_t>(gxc + X0 + c0)) = word;
          if (rr == 7) *reinterpret_cast<uint32_t*>(&F.Ac[pl][X0 + c0]) = word;
          if (c0 == 4) F.Lcc[pl][Y0 + rr] = last;
          if (rr == 7 && c0 == 4) F.BRc[pl][li] = last;
        } else {
          *reinterpret_cast<uint32_t*>(rec.y + static_cast<size_t>(gyl + Y0 + rr) * static_cast<size_t>(rec.pw) +
                                       static_cast<size_t>(gxl + X0 + c0)) = word;
          if (rr == n - 1) *reinterpret_cast<uint32_t*>(&F.A[X0 + c0]) = word;
          if (c0 + 4 == n) F.Lc[Y0 + rr] = last;
          if (((Y0 + rr) & 3) == 3) F.BR[((Y0 + rr) >> 2) * (tw * 4) + ((X0 + c0) >> 2)] = last;
        }
      }
      cell_block_sync();  // the next level reads this one's frontier entries (and the ref space is reused)
    }
  }
}

// One workgroup per tile.  Macroblock (x, y) depends on (x-1, y), (x-1, y-1),
// (x, y-1) and (x+1, y-1), so step s = x + 2y (tile-relative) may run all its
// macroblocks at once.  Only intra macroblocks are reconstructed here (the
// others were written by the inter kernel before), so a first pass over the
// tile's records marks the steps that hold one and the others are skipped.
// The workgroup has blockDim.x / 64 wavefronts (up to kIntraMaxWaves); its
// dynamic LDS holds one IntraScratch per wavefront and, when the launch
// carries key frames (with_front), the Frontier in front of them.
constexpr int kIntraMaxWaves = 16;
constexpr int kMaxStepWords = 16;  // steps tracked individually (512)
// Inter frames: up to one wavefront's worth of intra macroblocks per tile
// (in tiles of at most kFrontMax x kFrontMax macroblocks) are listed with
// their dependency level.
// The frontier entries of a macroblock that is already final in the frame
// (an inter frame's skip or inter macroblock): its bottom row, right column
// and bottom-right corner, as intra_mb leaves them for an intra one.
__device__ __forceinline__ void front_from_frame(Frontier& fr, const Planes rec, const TileDesc& td, int mbx, int mby,
                                                 int lane) {
  const int rx = mbx - td.x0, ry = mby - td.y0, X = mbx * 16, Y = mby * 16;
  const int g = lane >> 4, i = lane & 15;  // g 0: luma bottom, 1: luma right, 2: Cb, 3: Cr
  if (g < 2) {
    const uint8_t v = g == 0 ? rec.y[static_cast<size_t>(Y + 15) * rec.pw + X + i]
                             : rec.y[static_cast<size_t>(Y + i) * rec.pw + X + 15];
    if (g == 0) fr.bot[rx][i] = v;
    else fr.right[ry][i] = v;
    if (g == 0 && i == 15) fr.corner[(ry + 1) & 1][rx + 1][0] = v;
  } else {
    const int pl = g - 2, k = i & 7;
    const uint8_t* cp = pl ? rec.cr : rec.cb;
    const int CX = mbx * 8, CY = mby * 8;
    const uint8_t v = i < 8 ? cp[static_cast<size_t>(CY + 7) * rec.cw + CX + k]
                            : cp[static_cast<size_t>(CY + k) * rec.cw + CX + 7];
    if (i < 8) fr.bot[rx][16 + 8 * pl + k] = v;
    else fr.right[ry][16 + 8 * pl + k] = v;
    if (i == 7) fr.corner[(ry + 1) & 1][rx + 1][1 + pl] = v;
  }
  wave_sync();
}

constexpr int kMaxIntraList = 64;
// Polls of a band's published count before its wavefront gives up (about
// 0.1 s: a band waits at most for the bands above to fill their pipeline).
constexpr uint32_t kBandSpinMax = 1u << 16;
constexpr int kIntraBitWords = kFrontMax * kFrontMax / 32 + 1;
__device__ __forceinline__ void recon_intra_tile(const uint8_t* __restrict__ desc, const MBRec* __restrict__ recs,
                                                 const int16_t* __restrict__ res,
                                                 const TileStatus* __restrict__ status, const Planes rec, const int t,
                                                 const bool with_front, const bool with_cells, uint32_t* sync,
                                                 const int band_rows) {
  extern __shared__ __attribute__((aligned(16))) uint8_t intra_dyn[];
  __shared__ uint32_t step_mask[kMaxStepWords];
  __shared__ uint32_t s_bits[kIntraBitWords];  // intra macroblocks of the tile (local index)
  __shared__ uint32_t s_list[kMaxIntraList];   // local index | level << 16, raster order
  __shared__ uint32_t s_count[2];              // list length, highest level
  Frontier& front = *reinterpret_cast<Frontier*>(intra_dyn);
  IntraScratch* scratch = reinterpret_cast<IntraScratch*>(intra_dyn + (with_front ? sizeof(Frontier) : 0));
  const FrameHdr* hdr = reinterpret_cast<const FrameHdr*>(desc);
  const TileDesc* tiles = reinterpret_cast<const TileDesc*>(desc + desc_tiles_off());
  const int lane = lane_id(), wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const TileDesc td = tiles[t];
  const TileStatus st = status[t];
  if (st.n_intra == 0) return;
  const int tw = td.x1 - td.x0, th = td.y1 - td.y0;
  const int steps = (tw - 1) + 2 * (th - 1) + 1;
  const int mbc = hdr->mbc;
  // Key-frame tiles and tiles dense with intra macroblocks: by block levels.
  if (with_cells && cell_tile_fits(tw, th, nwaves) &&
      (hdr->inter == 0 || st.n_intra > static_cast<uint32_t>(kMaxIntraList))) {
    intra_cells(intra_dyn, recs, res, td, mbc, rec, st.mbs_done, lane, wave, nwaves);
    return;
  }
  // A key frame's tile, and an inter frame's tile that is mostly intra,
  // keeps its context in LDS (the inter macroblocks' edges come from the
  // frame, where the inter launch left them final).
  const bool keyf = with_front && (hdr->inter == 0 || 2 * st.n_intra > st.mbs_done) && tw <= kFrontMax &&
                    th <= kFrontMax;
  // A tile given several items (key_bands): each workgroup claims one of
  // them; a tile that takes another path runs it in the first only.
  const int nb = key_bands(tw, th, band_rows);
  int band = 0;
  if (nb > 1) {
    __shared__ uint32_t s_band;
    if (threadIdx.x == 0) s_band = atomicAdd(&sync[td.y0 * mbc + td.x0], 1u);
    __syncthreads();
    band = static_cast<int>(s_band);
    if (band >= nb || (!keyf && band > 0)) return;
  }
  // Otherwise, a tile with few intra macroblocks runs them by dependency
  // level: level 0 has no intra neighbour among left, above-left, above and
  // above-right, level k + 1 has one of level k at most.  Macroblocks of one
  // level neither read nor write each other's pixels, so a level is one
  // parallel round, usually one or two rounds against one per anti-diagonal.
  const bool levels = !keyf && tw <= kFrontMax && th <= kFrontMax && st.n_intra <= kMaxIntraList;
  const bool tracked = !keyf && !levels && steps <= 32 * kMaxStepWords;
  if (threadIdx.x < kMaxStepWords) step_mask[threadIdx.x] = 0;
  if (levels)
    for (int i = threadIdx.x; i < (tw * th + 31) / 32; i += blockDim.x) s_bits[i] = 0;
  __syncthreads();  // (orders the clears above)
  if (levels) {
    for (uint32_t i = threadIdx.x; i < st.mbs_done; i += blockDim.x) {
      const int ry = static_cast<int>(i) / tw, rx = static_cast<int>(i) - ry * tw;
      if (recs[(td.y0 + ry) * mbc + td.x0 + rx].kind == 2) atomicOr(&s_bits[i >> 5], 1u << (i & 31));
    }
    __syncthreads();
    if (wave == 0) {
      // The list in raster order, lane k = the k-th intra macroblock.
      uint32_t mine = 0xFFFFFFFFu, n = 0;
      const int nwords = (tw * th + 31) / 32;
      for (int w0 = 0; w0 < nwords; w0 += 64) {
        const uint32_t word = w0 + lane < nwords ? s_bits[w0 + lane] : 0u;
        const int cnt = __builtin_popcount(word);
        int pre = cnt;  // inclusive prefix of the counts over the lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const int o = __shfl_up(pre, d);
          if (lane >= d) pre += o;
        }
        const int base = static_cast<int>(n) + pre - cnt;
        // each lane's set bits go to lanes base .. base + cnt - 1
        uint32_t w = word;
        for (int j = 0; w; j++) {
          const uint32_t b = static_cast<uint32_t>(__builtin_ctz(w));
          w &= w - 1;
          const int dst = base + j;
          if (dst < kMaxIntraList) s_list[dst] = static_cast<uint32_t>((w0 + lane) * 32) + b;
        }
        n += static_cast<uint32_t>(__shfl(pre, 63));
      }
      wave_sync();
      if (n > static_cast<uint32_t>(kMaxIntraList)) n = 0xFFFFFFFFu;  // more than the tile status said: steps
      if (static_cast<uint32_t>(lane) < n) mine = s_list[lane];
      // Levels in raster order: a macroblock's intra neighbours come earlier.
      uint32_t lvl = 0, maxl = 0;
      for (uint32_t k = 0; n != 0xFFFFFFFFu && k < n; k++) {
        const int loc = static_cast<int>(__builtin_amdgcn_readlane(mine, static_cast<int>(k)));
        const int ry = loc / tw, rx = loc - ry * tw;
        uint32_t l = 0;
        auto dep = [&](bool present, int nloc) {  // an intra neighbour's level + 1
          if (!present) return;
          const uint64_t hit = __ballot(mine == static_cast<uint32_t>(nloc));
          if (hit) {
            const uint32_t v = __builtin_amdgcn_readlane(lvl, static_cast<int>(__builtin_ctzll(hit))) + 1u;
            l = v > l ? v : l;
          }
        };
        dep(rx > 0, loc - 1);                       // left
        dep(rx > 0 && ry > 0, loc - tw - 1);        // above-left
        dep(ry > 0, loc - tw);                      // above
        dep(ry > 0 && rx + 1 < tw, loc - tw + 1);   // above-right
        if (static_cast<uint32_t>(lane) == k) lvl = l;
        maxl = l > maxl ? l : maxl;
      }
      if (n != 0xFFFFFFFFu && static_cast<uint32_t>(lane) < n) s_list[lane] = mine | (lvl << 16);
      if (lane == 0) {
        s_count[0] = n;
        s_count[1] = maxl;
      }
    }
    __syncthreads();
    const uint32_t n = s_count[0], maxl = s_count[1];
    for (uint32_t lv = 0; n != 0xFFFFFFFFu && lv <= maxl; lv++) {
      for (uint32_t k = wave; k < n; k += nwaves) {
        const uint32_t e = s_list[k];
        if ((e >> 16) != lv) continue;
        const int local = static_cast<int>(e & 0xFFFFu);
        const int ry = local / tw, rx = local - ry * tw;
        const int mbx = td.x0 + rx, mby = td.y0 + ry;
        intra_mb(scratch[wave], res + static_cast<size_t>(mby * mbc + mbx) * kResPerMB, recs[mby * mbc + mbx], td, rec,
                 mbx, mby, nullptr, lane);
      }
      // the next level reads this one's pixels from the frame (see below)
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if (n != 0xFFFFFFFFu) return;
    // (otherwise every anti-diagonal below, untracked)
  }
  if (keyf) {
    // A key frame's tile as a dataflow over its macroblock rows instead of a
    // barrier per anti-diagonal: wavefront w owns rows w, w + W, ... and
    // visits their macroblocks in anti-diagonal order (x + 2y); macroblock
    // (x, y) waits only until row y - 1 has completed through column x + 1
    // (its above and above-right neighbours; left and above-left are older in
    // its own row and the row above), read from a per-row counter in LDS.
    // A slow macroblock (16 serial 4x4 blocks) then delays its dependants
    // only, not every macroblock of its anti-diagonal.  Every wait is for a
    // macroblock of a lower anti-diagonal, and each wavefront goes through
    // its macroblocks in that order, so every wait ends.  The frontier
    // entries a macroblock reads were written by the macroblocks it waits
    // for, and the ones it overwrites are read only by macroblocks that
    // complete before it (its own row, or the row below's left part, which
    // it waits for through row y - 1 when it is itself in row y + 1).
    //
    // A tall tile (key_bands > 1) is split into bands of band_rows rows, one
    // workgroup each; each workgroup claims the next band of its tile from
    // the tile's counter, so a band's workgroup only ever waits for one that
    // runs already (it claimed earlier).  A band's first row takes the row
    // above from the frame, waiting for the band above to have published
    // its last row through column x + 1: that row stores its bottom samples
    // device-coherently, waits for them, then moves the published count.
    // The wait is bounded: a count that never moved (which the zeroed
    // counters and the item count exclude) would leave wrong samples in
    // the band's rows rather than hold the launch.
    __shared__ int s_rowdone[kFrontMax];
    int by0 = 0, by1 = th;
    if (nb > 1) {
      by0 = band * band_rows;
      by1 = by0 + band_rows < th ? by0 + band_rows : th;
    }
    const uint32_t* prog_in = nb > 1 && by0 > 0 ? &sync[(td.y0 + by0) * mbc + td.x0] : nullptr;
    uint32_t* prog_out = nb > 1 && by1 < th ? &sync[(td.y0 + by1) * mbc + td.x0] : nullptr;
    for (int i = threadIdx.x; i < th; i += blockDim.x) s_rowdone[i] = 0;
    __syncthreads();
    volatile int* rowdone = s_rowdone;
    bool lost = false;  // this wavefront gave up on a published count
    const int bsteps = (tw - 1) + 2 * (by1 - by0 - 1) + 1;
    // The wavefront's macroblocks in the order above: step s, then its rows
    // ry = by0 + wave + k * nwaves with rx = s - 2 * (ry - by0) in the tile.
    // The next one's record and residual are loaded while this one runs.
    auto advance = [&](int& s, int& ry) -> bool {
      if (by0 + wave >= by1) return false;  // (a wavefront with no row in the band)
      for (;;) {
        ry += nwaves;
        if (ry >= by1) {
          ry = by0 + wave;
          if (++s >= bsteps) return false;
        }
        const int rx = s - 2 * (ry - by0);
        if (rx >= 0 && rx < tw) return true;
      }
    };
    uint32_t vrec = 0, w0 = 0, w1 = 0, w2 = 0;
    auto fetch = [&](int s, int ry) {
      const int mb = (td.y0 + ry) * mbc + td.x0 + s - 2 * (ry - by0);
      const uint32_t* rp = reinterpret_cast<const uint32_t*>(recs + mb);
      vrec = lane < 12 ? rp[lane] : 0u;
      const uint32_t* src = reinterpret_cast<const uint32_t*>(res + static_cast<size_t>(mb) * kResPerMB);
      w0 = src[lane];
      w1 = src[lane + 64];
      w2 = src[lane + 128];
    };
    int s = 0, ry = by0 + wave - nwaves;
    bool have = advance(s, ry);
    if (have) fetch(s, ry);
    while (have) {
      const int rx = s - 2 * (ry - by0);
      const uint32_t cr = vrec, c0 = w0, c1 = w1, c2 = w2;
      // (the lane index made opaque per macroblock: otherwise the compiler
      // hoists every lane-dependent term of the macroblock code out of the
      // loop and keeps them all live, past the registers the launch has)
      int ln = lane;
      asm volatile("" : "+v"(ln));
      int s2 = s, ry2 = ry;
      const bool have2 = advance(s2, ry2);
      if (have2) fetch(s2, ry2);
      const int need = rx + 2 < tw ? rx + 2 : tw;
      int edge = 0;
      if (ry > by0) {
        while (rowdone[ry - 1] < need) __builtin_amdgcn_s_sleep(1);
        asm volatile("" ::: "memory");
      } else if (prog_in) {
        uint32_t spins = 0;
        while (!lost && __hip_atomic_load(prog_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                            static_cast<uint32_t>(need)) {
          __builtin_amdgcn_s_sleep(2);
          if (++spins >= kBandSpinMax) lost = true;
        }
        asm volatile("" ::: "memory");
        edge = kEdgeIn;
      }
      if (ry == by1 - 1 && prog_out) edge |= kEdgeOut;
      const uint32_t local = static_cast<uint32_t>(ry * tw + rx);
      if (local < st.mbs_done) {
        const MBRec r = rec_of_lanes(cr);
        if (r.kind == 2) {
          const bool z = r.coef_count == 0;  // (then the residual words were never written)
          if (r.tx == 2 && (r.wmask & 0x30001u) == 0x30001u)
            intra_mb16(scratch[wave], z ? 0u : c0, z ? 0u : c1, z ? 0u : c2, r, td, rec, td.x0 + rx, td.y0 + ry,
                       front, ln, edge);
          else
            intra_mb_w(scratch[wave], z ? 0u : c0, z ? 0u : c1, z ? 0u : c2, r, td, rec, td.x0 + rx, td.y0 + ry,
                       &front, ln, edge);
        } else {
          front_from_frame(front, rec, td, td.x0 + rx, td.y0 + ry, ln);
        }
      }
      if (edge & kEdgeOut) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (lane == 0) __hip_atomic_store(prog_out, static_cast<uint32_t>(rx + 1), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
      }
      // the frontier writes land before the row's count moves on
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      rowdone[ry] = rx + 1;
      s = s2;
      ry = ry2;
      have = have2;
    }
    return;
  }
  if (tracked) {
    for (uint32_t i = threadIdx.x; i < st.mbs_done; i += blockDim.x) {
      const int ry = static_cast<int>(i) / tw, rx = static_cast<int>(i) - ry * tw;
      if (recs[(td.y0 + ry) * mbc + td.x0 + rx].kind == 2) {
        const int s = rx + 2 * ry;
        atomicOr(&step_mask[s >> 5], 1u << (s & 31));
      }
    }
  }
  __syncthreads();
  for (int s = 0; s < steps; s++) {
    if (tracked && !((step_mask[s >> 5] >> (s & 31)) & 1u)) continue;  // uniform across the workgroup
    // Macroblocks on this step: rows ry with rx = s - 2*ry in [0, tw).
    int ry_lo = s - (tw - 1) > 0 ? (s - (tw - 1) + 1) / 2 : 0;
    int ry_hi = s / 2 < th - 1 ? s / 2 : th - 1;
    for (int ry = ry_lo + wave; ry <= ry_hi; ry += nwaves) {
      int rx = s - 2 * ry;
      uint32_t local = static_cast<uint32_t>(ry * tw + rx);
      if (local >= st.mbs_done) continue;
      int mbx = td.x0 + rx, mby = td.y0 + ry;
      const MBRec r = recs[mby * mbc + mbx];
      if (r.kind != 2) continue;
      intra_mb(scratch[wave], res + static_cast<size_t>(mby * mbc + mbx) * kResPerMB, r, td, rec, mbx, mby,
               keyf ? &front : nullptr, lane);
    }
    // Macroblocks of the next step read what other wavefronts of this
    // workgroup stored in this one.  In a key frame that is the LDS frontier
    // and the barrier suffices.  Otherwise they read the frame: the barrier
    // orders the stores, but the CU's vector L1 may still hold lines fetched
    // before them (another wavefront's neighbour loads touch the same
    // 128-byte lines), so the acquire at agent scope invalidates L1 before
    // the next step's loads.
    __syncthreads();
    if (!keyf) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
}

__global__ __launch_bounds__(64 * kIntraMaxWaves) void ebiv_recon_intra(const FrameJob* __restrict__ jobs,
                                                                          const WorkItem* __restrict__ items,
                                                                          int flags) {
  const WorkItem it = item_of(items);
  ConstJob& job = job_of(jobs, it.job);
  recon_intra_tile(job.desc, job.recs, job.res, job.status, planes_of(job.rec), static_cast<int>(it.idx),
                   (flags & 1) != 0, (flags & 2) != 0, band_sync(job.res, job.mbc, job.mbr), flags >> 8);
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

// Eight pixels of one row per thread: Y as one 8-byte load and Cb, Cr as one
// dword each (the decoder's padded planes: rows of a multiple of 16 luma and
// 8 chroma bytes, so these loads are aligned and stay inside the plane), and
// two 16-byte stores where the destination row allows, else dword or byte
// stores.  HBM-bound: 1.5 bytes read and 4 written per pixel.
__device__ __forceinline__ void rgba_span(const uint8_t* __restrict__ y, int ys, const uint8_t* __restrict__ cb,
                                          const uint8_t* __restrict__ cr, int cs, int w, int h, uint8_t* __restrict__ dst,
                                          int64_t pitch, uint32_t idx) {
  const uint32_t q8 = static_cast<uint32_t>(w + 7) >> 3;
  if (idx >= q8 * static_cast<uint32_t>(h)) return;
  const uint32_t row = idx / q8;
  const int col = static_cast<int>(idx - row * q8) * 8;
  const uint2 yv = *reinterpret_cast<const uint2*>(y + static_cast<size_t>(row) * ys + col);
  const size_t co = static_cast<size_t>(row >> 1) * cs + (col >> 1);
  const uint32_t bv = *reinterpret_cast<const uint32_t*>(cb + co);
  const uint32_t rv = *reinterpret_cast<const uint32_t*>(cr + co);
  uint32_t px[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t yb = ((k < 4 ? yv.x : yv.y) >> (8 * (k & 3))) & 0xFFu;
    px[k] = rgba_px(static_cast<int32_t>(yb), static_cast<int32_t>((bv >> (8 * (k >> 1))) & 0xFFu),
                    static_cast<int32_t>((rv >> (8 * (k >> 1))) & 0xFFu));
  }
  uint8_t* o = dst + static_cast<int64_t>(row) * pitch + static_cast<int64_t>(col) * 4;
  const int n = w - col < 8 ? w - col : 8;
  if (n == 8 && (reinterpret_cast<uintptr_t>(o) & 15) == 0) {
    reinterpret_cast<uint4*>(o)[0] = make_uint4(px[0], px[1], px[2], px[3]);
    reinterpret_cast<uint4*>(o)[1] = make_uint4(px[4], px[5], px[6], px[7]);
  } else if ((reinterpret_cast<uintptr_t>(o) & 3) == 0) {
    for (int k = 0; k < n; k++) reinterpret_cast<uint32_t*>(o)[k] = px[k];
  } else {
    for (int k = 0; k < n; k++)
      for (int b = 0; b < 4; b++) o[4 * k + b] = static_cast<uint8_t>(px[k] >> (8 * b));
  }
}

// One frame, its parameters as kernel arguments.
__global__ __launch_bounds__(256) void ebiv_rgba(RgbaJob j) {
  rgba_span(j.y, j.ys, j.cb, j.cr, j.cs, j.w, j.h, j.dst, j.pitch, blockIdx.x * 256u + threadIdx.x);
}

// Many frames in one launch (the batch step's presentation): blockIdx.y is
// the frame, read from a job table through the constant address space.
using ConstRgbaJob = __attribute__((address_space(4))) const RgbaJob;
__global__ __launch_bounds__(256) void ebiv_rgba_frames(const RgbaJob* __restrict__ jobs) {
  const ConstRgbaJob& j = ((ConstRgbaJob*)jobs)[blockIdx.y];
  rgba_span(j.y, j.ys, j.cb, j.cr, j.cs, j.w, j.h, j.dst, j.pitch, blockIdx.x * 256u + threadIdx.x);
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
  size_t lds = kRecStage * sizeof(MBRec) + 2 * static_cast<size_t>(max_tile_w) * sizeof(uint32_t);
  hipLaunchKernelGGL(ebiv_parse, dim3(nitems), dim3(64), lds, s, jobs, items);
}

void launch_recon_inter(hipStream_t s, int ngroups, const FrameJob* jobs, const WorkItem* items, int njobs) {
  hipLaunchKernelGGL(ebiv_recon_inter, dim3(ngroups), dim3(256), 0, s, jobs, items, njobs);
}

// waves: wavefronts per tile (1-16); key: the items include key-frame tiles,
// which then use the LDS frontier.
// cells: key-frame tiles and tiles dense with intra macroblocks take the
// block-level path (its LDS overlays the other paths').
// band_rows: key_bands of the key-frame tiles (the items hold one per band).
void launch_recon_intra(hipStream_t s, int nitems, const FrameJob* jobs, const WorkItem* items, int waves, bool key,
                        bool cells, int band_rows) {
  waves = waves < 1 ? 1 : (waves > kIntraMaxWaves ? kIntraMaxWaves : waves);
  size_t lds = static_cast<size_t>(waves) * sizeof(IntraScratch) + (key ? sizeof(Frontier) : 0);
  if (cells && cell_lds_bytes(waves) > lds) lds = cell_lds_bytes(waves);
  band_rows = band_rows < 0 ? 0 : (band_rows > kFrontMax ? kFrontMax : band_rows);
  hipLaunchKernelGGL(ebiv_recon_intra, dim3(nitems), dim3(64 * waves), lds, s, jobs, items,
                     (key ? 1 : 0) | (cells ? 2 : 0) | (band_rows << 8));
}

void launch_rgba(hipStream_t s, const RgbaJob& j) {
  const uint32_t n = static_cast<uint32_t>((j.w + 7) / 8) * static_cast<uint32_t>(j.h);
  hipLaunchKernelGGL(ebiv_rgba, dim3((n + 255) / 256), dim3(256), 0, s, j);
}

// max_px8: the most 8-pixel spans of any frame in the table.
void launch_rgba_frames(hipStream_t s, const RgbaJob* jobs, int njobs, uint32_t max_px8) {
  hipLaunchKernelGGL(ebiv_rgba_frames, dim3((max_px8 + 255) / 256, njobs), dim3(256), 0, s, jobs);
}

}  // namespace ebiv
