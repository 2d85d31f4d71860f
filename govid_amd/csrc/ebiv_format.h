// ebiv_format.h — EBIV bitstream model shared by the host runtime, the
// encoder and (through plain constants) the device kernels.
//
// Everything here is data the format fixes: the context layout and alphabet
// sizes (ebiv/model.go:21-213), the partition rectangles (model.go:91-96), the
// luma interpolation taps (ebiv/motioncomp.go:30-35), the angular intra modes
// (ebiv/predict.go:36-43), and the Go varint rules (encoding/binary).
#pragma once

#include <cstddef>
#include <cstdint>

namespace ebiv {

constexpr int kMB = 16;      // luma macroblock size
constexpr int kCMB = 8;      // chroma block per macroblock (4:2:0)
constexpr int kDCPred = 128; // intra prediction with no neighbours
constexpr int kMaxQP = 63;
constexpr int kSDHMinSpan = 4;      // residual.go:21-25
constexpr int kMVClamp = 4096;      // motioncomp.go:19
constexpr int kRansScaleBits = 12;  // rans.go:20-25
constexpr uint32_t kRansM = 1u << kRansScaleBits;
constexpr uint32_t kRansLower = 1u << 23;
constexpr int kRansStates = 4;
constexpr long kMaxDecodePixels = 8192L * 4320L; // codec.go:123

enum IntraMode { kModeDC, kModeV, kModeH, kModeTM, kModeD45, kModeD27, kModeD23, kModeD135, kModeD117, kModeD113, kNumIntraModes };
enum MBMode { kMBSkip, kMBInter, kMBIntra };
enum RefSel { kRefLast, kRefGolden };
enum CBPBits { kCBPLuma = 1, kCBPCb = 2, kCBPCr = 4 };
enum PartMode { kPartWhole, kPart16x8, kPart8x16, kPart8x8 };
enum Token { kTZero, kTOne, kTTwo, kTThree, kTFour, kTEscape, kTEOB };

enum Ctx {
  kCtxTxSize, kCtxLumaMode, kCtxChromaMode, kCtxMBMode, kCtxRef, kCtxPart, kCtxCBP,
  kCtxSign, kCtxEscClass, kCtxMVClassX, kCtxMVClassY, kCtxBypass, kCtxTokenBase
};
constexpr int kNumBands = 5;
constexpr int kNumPrevClass = 3;
constexpr int kLumaTokenCtx = 3 * kNumBands * kNumPrevClass; // 45
constexpr int kChromaTokenCtx = kNumBands * kNumPrevClass;   // 15
constexpr int kNumContexts = kCtxTokenBase + kLumaTokenCtx + kChromaTokenCtx; // 72
constexpr int kMaxAlphabet = 16;

inline int alphabet_size(int c) {
  constexpr int a[kCtxTokenBase] = {3, 10, 10, 3, 2, 4, 8, 2, 16, 16, 16, 2};
  return c < kCtxTokenBase ? a[c] : 7;
}
inline int token_ctx(int plane, int tx, int band, int prev) {
  return plane == 0 ? kCtxTokenBase + (tx * kNumBands + band) * kNumPrevClass + prev
                    : kCtxTokenBase + kLumaTokenCtx + band * kNumPrevClass + prev;
}
inline int band_of(int i) { return i == 0 ? 0 : i <= 2 ? 1 : i <= 8 ? 2 : i <= 20 ? 3 : 4; }
inline int class_for_mag(int m) { return m <= 0 ? 0 : m == 1 ? 1 : 2; }

constexpr int kPartCount[4] = {1, 2, 2, 4};
constexpr int kPartRect[4][4][4] = {
    {{0, 0, 16, 16}},
    {{0, 0, 16, 8}, {0, 8, 16, 8}},
    {{0, 0, 8, 16}, {8, 0, 8, 16}},
    {{0, 0, 8, 8}, {8, 0, 8, 8}, {0, 8, 8, 8}, {8, 8, 8, 8}},
};
constexpr int32_t kLumaTaps[4][4] = {{0, 64, 0, 0}, {-6, 55, 18, -3}, {-8, 40, 40, -8}, {-3, 18, 55, -6}};

struct AngMode { int angle, inv_angle; bool vertical; };
constexpr AngMode kAngModes[kNumIntraModes] = {
    {0, 0, false}, {0, 0, false}, {0, 0, false}, {0, 0, false},
    {32, 0, true}, {13, 0, true}, {-13, -630, true},
    {32, 0, false}, {13, 0, false}, {-13, -630, false},
};

// Status codes of the decoder, one per Go sentinel error class
// (ebiv/format.go:110-119).
enum Status { kOK = 0, kErrCorrupt = 1, kErrUnsupported = 2, kErrNoKeyframe = 3, kErrDimensions = 4 };

// encoding/binary.Uvarint: > 0 bytes read, 0 truncated, < 0 overflow.
inline int uvarint(const uint8_t* b, size_t len, uint64_t* out) {
  uint64_t x = 0;
  unsigned s = 0;
  for (size_t i = 0; i < len; i++) {
    if (i == 10) return -static_cast<int>(i + 1);
    uint8_t c = b[i];
    if (c < 0x80) {
      if (i == 9 && c > 1) return -static_cast<int>(i + 1);
      *out = x | static_cast<uint64_t>(c) << s;
      return static_cast<int>(i) + 1;
    }
    x |= static_cast<uint64_t>(c & 0x7f) << s;
    s += 7;
  }
  *out = 0;
  return 0;
}
inline int varint(const uint8_t* b, size_t len, int64_t* out) {
  uint64_t ux = 0;
  int n = uvarint(b, len, &ux);
  int64_t x = static_cast<int64_t>(ux >> 1);
  if (ux & 1) x = ~x;
  *out = x;
  return n;
}
inline size_t put_uvarint(uint8_t* b, uint64_t v) {
  size_t i = 0;
  while (v >= 0x80) { b[i++] = static_cast<uint8_t>(v) | 0x80; v >>= 7; }
  b[i++] = static_cast<uint8_t>(v);
  return i;
}
inline size_t put_varint(uint8_t* b, int64_t v) {
  uint64_t ux = static_cast<uint64_t>(v) << 1;
  if (v < 0) ux = ~ux;
  return put_uvarint(b, ux);
}

inline int align_up(int v, int a) { return (v + a - 1) & ~(a - 1); }

}  // namespace ebiv
