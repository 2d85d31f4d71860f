// ebiv_device.h — data layouts shared by the host runtime and the gfx950
// kernels.  Plain structs and constexpr helpers only (constexpr functions are
// callable from device code), so host-only code builds against it too.
#pragma once

#include <cstddef>
#include <cstdint>

namespace ebiv {

// One entropy-coded tile of a coded frame (ebiv/frame.go:162-177).
struct TileDesc {
  uint32_t byte_off;   // tile stream offset inside the frame's device packet
  uint32_t byte_len;   // tile stream length
  uint16_t x0, x1;     // macroblock column span [x0, x1)
  uint16_t y0, y1;     // macroblock row span [y0, y1)
  uint32_t coef_base;  // first coefficient slot owned by this tile
  uint32_t pad;
};
static_assert(sizeof(TileDesc) == 24, "TileDesc layout");

// The frame's 72 static rANS tables (rans.go:120-148) as the entropy kernel
// holds them in registers: 15 vectors of 64 lanes, one packed entry per lane,
//   entry = start << 16 | freq   (start <= 4096, freq <= 4096),
// kNoSym past a context's alphabet and on every lane of an unused context.
//   vector 0: sign (lanes 0-15), MB mode (16-31), transform size (32-47), CBP (48-63)
//   vector 1: luma mode, chroma mode, reference, partition (16 lanes each)
//   vector 2: escape class, MV class x, MV class y (16 lanes each)
//   vector 3 + 3*set + band/2, set = luma transform index 0-2 or 3 for chroma:
//     the token context (set, band, prev) at lanes 8*((band & 1)*3 + prev) + symbol
// The bypass context is fixed at {2048, 2048} (rans.go:312-322) and has no entry.
constexpr int kNumTableVecs = 15;
constexpr uint32_t kNoSym = 0xFFFFFFFFu;
// Where the sign context's table starts in the table vectors (vector 0, the
// first lane group): a frame without one takes the generic grammar decoder.
constexpr int kSignVec = 0, kSignLane = 0;
// Vector and first lane of a context's entries; false for the bypass context.
constexpr bool table_place(int ctx, int* vec, int* lane0) {
  // contexts 0-11 in enum order: tx, luma mode, chroma mode, MB mode, ref,
  // part, CBP, sign, escape class, MV x, MV y, bypass
  constexpr int8_t kVec[12] = {0, 1, 1, 0, 1, 1, 0, 0, 2, 2, 2, -1};
  constexpr int8_t kLane[12] = {32, 0, 16, 16, 32, 48, 48, 0, 0, 16, 32, -1};
  if (ctx < 12) {
    *vec = kVec[ctx];
    *lane0 = kLane[ctx];
    return ctx != 11;
  }
  int t = ctx - 12, set = 0, band = 0, prev = 0;
  if (t < 45) { set = t / 15; band = (t % 15) / 3; prev = t % 3; }
  else { set = 3; band = (t - 45) / 3; prev = (t - 45) % 3; }
  *vec = 3 + 3 * set + band / 2;
  *lane0 = 8 * ((band & 1) * 3 + prev);
  return true;
}

// Per-frame parameters: the coded header (frame.go:181-220), quantiser
// (quant.go:38-51) and the frame's symbol tables in kernel layout (above).
struct FrameHdr {
  int32_t mbc, mbr;       // macroblock grid of the padded frame
  int32_t ntiles;
  int32_t inter;          // 1: decodeInterMB grammar, 0: intra grammar
  int32_t qdc, qac;       // dequantisation steps
  int32_t grid_cols, grid_rows;
  uint32_t used[3];       // bit c: context c has a usable table
  int32_t max_tile_w;     // widest tile in macroblocks (MV row buffers)
  uint32_t vt[kNumTableVecs][64];
};

// Per-frame device descriptor = FrameHdr, TileDesc[ntiles], uint16 col_tile[mbc],
// uint16 row_tile[mbr], laid out back to back (each part 16-byte aligned).
constexpr size_t align16(size_t v) { return (v + 15) & ~size_t(15); }
constexpr size_t desc_tiles_off() { return align16(sizeof(FrameHdr)); }
constexpr size_t desc_coltile_off(int ntiles) { return align16(desc_tiles_off() + sizeof(TileDesc) * ntiles); }
constexpr size_t desc_rowtile_off(int ntiles, int mbc) { return align16(desc_coltile_off(ntiles) + 2 * size_t(mbc)); }
constexpr size_t desc_size(int ntiles, int mbc, int mbr) { return align16(desc_rowtile_off(ntiles, mbc) + 2 * size_t(mbr)); }

// Macroblock record written by the entropy kernel and read by the
// reconstruction kernels.  kind: 0 skip, 1 inter, 2 intra (decode_inter.go:8-62),
// kKindUnreached for a macroblock after the one on which its tile's stream
// failed (no pixel of it is written).
constexpr uint32_t kKindUnreached = 3;
struct alignas(16) MBRec {
  uint8_t kind, part, cbp, tx;
  uint8_t refsel, cmode, pad0, pad1;
  // Which outputs Go's decoder writes for this macroblock: bits 0-15 luma
  // blocks in decode order (all set for an uncoded inter luma), bit 16 Cb,
  // bit 17 Cr.  A fully decoded macroblock has every bit it uses set; only the
  // macroblock on which a tile's stream failed carries a partial mask.
  uint32_t wmask;
  // Coefficient entries: scan index (bits 0-7) | block << 8 | level << 13,
  // the block 0-15 luma in decode order, 16 Cb, 17 Cr.
  uint32_t coef_off;    // first coefficient entry of this macroblock
  uint32_t coef_count;  // number of coefficient entries
  int16_t mv[4][2];     // partition vectors (quarter-pel), skip/whole use mv[0]
  uint32_t lmode[2];    // intra luma modes, 4 bits each, in block decode order
  uint32_t pad2;
};
static_assert(sizeof(MBRec) == 48, "MBRec layout");

// Coefficient entry: bits 0-7 position (row-major in its block), bits 8-12
// block id (0-15 luma in decode order, 16 Cb, 17 Cr), bits 13-31 the signed
// level (|level| <= 32772 always fits 19 bits).
constexpr int kCoefPerMB = 384;
// Residual of an intra macroblock, formed by the inter-reconstruction launch
// for the intra launch: luma 256 then Cb 64, Cr 64, each value clamped to
// int16 (the reconstruction clamps pred + residual to [0, 255] with pred in
// [0, 255], so clamping the residual to [-32768, 32767] changes no sample).
constexpr int kResPerMB = 384;

// Tiles of at most kFrontMax x kFrontMax macroblocks whose macroblocks are
// all intra (key frames, and inter frames coded all intra) reconstruct as a
// dataflow over their macroblock rows with the context in LDS.  With
// band_rows > 0 a taller tile is split into bands of band_rows rows, one
// workgroup each, so that a big tile's rows spread over many CUs; a band's
// first row then takes the row above from the frame, after the band above
// has published it.  key_bands is the number of workgroups (intra launch
// items) of such a tile; the host and the kernel both use it.  (An
// inter-frame tile that turns out not to be all intra runs in the first.)
constexpr int kFrontMax = 128;
constexpr int key_bands(int tw, int th, int band_rows) {
  return band_rows > 0 && th > band_rows && tw <= kFrontMax && th <= kFrontMax ? (th + band_rows - 1) / band_rows
                                                                                 : 1;
}
// The band counters of a payload: one uint32 per macroblock after its
// residuals (kResPerMB each).  For a key-frame tile at columns [x0, x1),
// rows [y0, y1): entry (y0, x0) counts the tile's claimed bands, entry
// (y, x0) for y > y0 the macroblocks of row y - 1 published to the band
// starting at row y.  The entropy kernel zeroes every tile's entries.
constexpr int kResSyncPerMB = 2;  // int16 slots per macroblock that the counters take

// Per-tile outcome of the entropy kernel.
struct TileStatus {
  int32_t err;        // 0 ok, else a Status code (always corrupt)
  uint32_t mbs_done;  // macroblocks (tile raster order) that carry a record
  uint32_t nsym;      // rANS decode calls, bypass bits included
  uint32_t n_intra;   // intra macroblocks among them
};

// One padded frame buffer on the device (ebiv/frame.go:19-37): luma pw x ph,
// chroma cw x ch, rows packed (stride = width).
struct Planes {
  uint8_t* y;
  uint8_t* cb;
  uint8_t* cr;
  int pw, ph;
  int cw, ch;
};

// One coded payload in a launch: where its packet, descriptor and
// entropy-kernel outputs live, and the frame buffers its reconstruction uses.
struct FrameJob {
  const uint8_t* pkt;
  const uint8_t* desc;
  MBRec* recs;
  uint32_t* coefs;
  TileStatus* status;
  int16_t* res;      // intra macroblocks' residuals, 384 per macroblock (see kResPerMB)
  int32_t mbc, mbr;  // the frame's macroblock grid
  Planes rec, ref, gold;
};

// Macroblocks per inter-reconstruction work item: two per wavefront of a
// 4-wavefront workgroup.
constexpr uint32_t kInterMbsPerWave = 2;
constexpr uint32_t kInterGroupMbs = 4 * kInterMbsPerWave;

// One workgroup's share of a launch: job index and a tile or macroblock index.
struct WorkItem {
  uint32_t job;
  uint32_t idx;
};

// One YCbCr 4:2:0 -> RGBA conversion (frame.go:8-67): a decoded frame's
// padded planes in, w x h RGBA pixels out at dst with `pitch` bytes per row.
struct RgbaJob {
  const uint8_t* y;
  const uint8_t* cb;
  const uint8_t* cr;
  uint8_t* dst;
  int64_t pitch;
  int32_t ys, cs, w, h;
};

// Error detail codes of TileStatus::err >> 8 (low byte is the Status).
enum TileErr { kTileOK = 0, kTileShort = 1, kTileUnused = 2, kTileUnderrun = 3 };

}  // namespace ebiv
