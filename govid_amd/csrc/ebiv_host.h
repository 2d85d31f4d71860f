// ebiv_host.h — the host half of Codec.Decode: frame header
// (ebiv/format.go:376-409), configure's checks (ebiv/codec.go:129-153), the
// golden-override split (codec.go:162-173), the coded header, table and tile
// framing (ebiv/frame.go:197-274, ebiv/rans.go:396-481), and Go's state
// transitions of the table history and the ring position.  Plain C++ with no
// device code: the runtime (ebiv_runtime.cpp) derives its sessions from
// HostCodec and adds the device frame pool; host-only builds (the sanitizer
// replay under tests/native) use it as it is.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ebiv_gpu.h"
#include "ebiv_device.h"
#include "ebiv_format.h"

namespace ebiv {

// prevFreqs slot: Go keeps the backing array across key-frame wipes
// (rans.go:397-405, 473-481), so a re-grown slice shows its old values.
struct PrevFreq {
  uint32_t v[kMaxAlphabet];
  int len = 0;
  bool alloc = false;
};

// One coded payload (a frame's own, or a golden override) after its framing
// and tables were parsed: the device descriptor (FrameHdr, TileDesc[], the
// column and row tile maps; ebiv_device.h) in host memory.
struct HostPayload {
  std::vector<uint8_t> desc;
  int ntiles = 0;
  int max_tile_w = 0;
  bool inter = false;
};

// One packet's host-side state.
struct HostFrame {
  int rc = EBIV_OK;  // host-side failure (nothing runs on the device)
  std::string msg;
  bool device_work = false;  // a coded payload or a raw copy runs on the device
  bool raw = false;
  bool key = false, inter = false, has_ov = false;
  int ring_slot = -1;
  int main_slot = -1, ov_slot = -1;  // payload slots (HostCodec::payloads)
  int ov_rc = EBIV_OK;               // host-side failure of the override payload
  std::string ov_msg;
  std::vector<PrevFreq> saved_prev;  // table history before the override's tables
  const uint8_t* host = nullptr;
  size_t len = 0;
  size_t body_off = 0;
  bool golden_written = false;       // set by resolve: golden took this frame's output
};

struct HostCodec {
  virtual ~HostCodec() = default;

  int ring_size = 3;
  int window = 1;
  bool ready = false;
  int W = 0, H = 0, CW = 0, CH = 0, mbc = 0, mbr = 0;
  int ystride = 0, cstride = 0;  // host ring image strides (geometry.go:22-30)
  int next = 0;                  // ring position (codec.go:84-88)
  PrevFreq prev[kNumContexts];
  std::vector<HostPayload> payloads;  // by payload slot
  std::string msg;

  int fail(int code, const std::string& m) {
    msg = m;
    return code;
  }

  // parseTablesInto (rans.go:396-471) into a frame header's table vectors,
  // updating prev exactly as Go does (including on failure).
  int parse_tables(const uint8_t* b, size_t len, bool key, FrameHdr& hdr, size_t* consumed);

  // decodeState.decodeCoded up to the tile fan-out (decode_intra.go:55-64,
  // frame.go:197-274): payload slot `si`'s descriptor.  body is a sub-range
  // of the packet starting at pkt_base (tile offsets are packet-relative).
  int prepare_coded(const uint8_t* pkt_base, const uint8_t* body, size_t len, bool inter, bool key, int si);

  // Codec.Decode up to the device work (codec.go:70-100, 162-173): header,
  // configure, ring slot, raw check, override split, payload preparation.
  // Payload slots are numbered from *next_slot.  Returns false when the
  // frame must end its session's window for the step: the override's tables
  // are parsed only if its key frame decodes.
  bool prepare_frame(HostFrame& f, int* next_slot);

  // Go's outcome of a prepared frame, given whether its payloads' tiles all
  // decoded (codec.go:174-190).  Restores the table history where Go never
  // reached the override's tables; sets f.golden_written when golden now
  // holds this frame's key or (partly) rewritten override.
  int resolve(HostFrame& f, bool main_ok, bool ov_ok);

 protected:
  // The device frame pool of the geometry configure just set (none here).
  virtual int allocate_frames() { return EBIV_OK; }
  // Device buffers of payload slot si (none here).
  virtual int reserve_payload(int) { return EBIV_OK; }

 private:
  int configure(int w, int h);
  bool prepare_frame_inner(HostFrame& f, int* next_slot);
};

}  // namespace ebiv
