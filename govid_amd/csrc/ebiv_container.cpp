// ebiv_container.cpp — the EBIV container reader on the host, behind the C ABI.
//
// Restates ebiv.NewDemuxer / Reader (ebiv/reader.go:32-258) and the file
// header and record formats (ebiv/format.go:151-328) over an in-memory copy
// of the file: a fixed 48-byte header, self-delimiting frame records (an
// 8-byte prefix opened by the 0xEB17 sync word), an index of 16-byte entries
// and a 20-byte footer holding the index's offset, count and CRC-32.  A
// missing or damaged footer falls back to scanning the records, as Go does.
// Integer widths follow Go's: offsets are uint64, index arithmetic int64,
// and timestamps the int64 nanoseconds of time.Duration (wrapping as Go's
// multiplication wraps).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/ebiv_gpu.h"

namespace {

constexpr size_t kFileHeaderSize = 48;
constexpr size_t kFrameRecordSize = 8;
constexpr size_t kIndexEntrySize = 16;
constexpr size_t kFooterSize = 20;
constexpr uint16_t kVersion = 4;
constexpr uint32_t kMaxFrameSize = 256u << 20;
constexpr uint16_t kFrameSync = 0xEB17;
constexpr uint32_t kFlagKeyframe = 1;
constexpr int64_t kSecond = 1000000000;

uint16_t le16(const uint8_t* p) { return static_cast<uint16_t>(p[0] | (p[1] << 8)); }
uint32_t le32(const uint8_t* p) {
  return static_cast<uint32_t>(p[0]) | (static_cast<uint32_t>(p[1]) << 8) | (static_cast<uint32_t>(p[2]) << 16) |
         (static_cast<uint32_t>(p[3]) << 24);
}
uint64_t le64(const uint8_t* p) { return static_cast<uint64_t>(le32(p)) | (static_cast<uint64_t>(le32(p + 4)) << 32); }

// hash/crc32.ChecksumIEEE: the reflected CRC-32 with polynomial 0xEDB88320.
uint32_t crc32_ieee(const uint8_t* p, size_t n) {
  static uint32_t table[256];
  static bool ready = false;
  if (!ready) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    ready = true;
  }
  uint32_t c = 0xFFFFFFFFu;
  for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

// Go's int64 multiply wraps; so does this one.
int64_t mul_wrap(int64_t a, int64_t b) {
  return static_cast<int64_t>(static_cast<uint64_t>(a) * static_cast<uint64_t>(b));
}

// frameTime (format.go:250-253).
int64_t frame_time(uint32_t num, uint32_t den, int64_t i) {
  return mul_wrap(mul_wrap(i, static_cast<int64_t>(den)), kSecond) / static_cast<int64_t>(num);
}

// frameIndexAt (format.go:263-275).
int64_t frame_index_at(uint32_t num, uint32_t den, int64_t t) {
  if (t <= 0) return 0;
  int64_t i = mul_wrap(t, static_cast<int64_t>(num)) / mul_wrap(static_cast<int64_t>(den), kSecond);
  while (frame_time(num, den, i + 1) <= t) i++;
  while (i > 0 && frame_time(num, den, i) > t) i--;
  return i;
}

struct Fail {
  int code;
  std::string msg;
};

// FileHeader.unmarshal (format.go:208-247).
bool parse_header(const uint8_t* b, size_t n, ebiv_file_header& h, Fail& f) {
  if (n < kFileHeaderSize) {
    f = {EBIV_ERR_CORRUPT, "ebiv: corrupt stream: file header truncated"};
    return false;
  }
  if (std::memcmp(b, "EBIV", 4) != 0) {
    f = {EBIV_ERR_BAD_MAGIC, "ebiv: not an EBIV file"};
    return false;
  }
  std::memset(&h, 0, sizeof(h));
  h.version = le16(b + 4);
  if (h.version != kVersion) {
    f = {EBIV_ERR_VERSION, "ebiv: unsupported format version: version " + std::to_string(h.version) + ", want " +
                               std::to_string(kVersion)};
    return false;
  }
  if (const uint16_t size = le16(b + 6); size != kFileHeaderSize) {
    f = {EBIV_ERR_CORRUPT, "ebiv: corrupt stream: header size " + std::to_string(size) + ", want 48"};
    return false;
  }
  const uint32_t w = le32(b + 8), ht = le32(b + 12);
  if (w == 0 || w > 65535 || ht == 0 || ht > 65535) {
    f = {EBIV_ERR_CORRUPT, "ebiv: corrupt stream: dimensions " + std::to_string(w) + "x" + std::to_string(ht) +
                               " out of range"};
    return false;
  }
  h.width = static_cast<int32_t>(w);
  h.height = static_cast<int32_t>(ht);
  h.fps_num = le32(b + 16);
  h.fps_den = le32(b + 20);
  if (h.fps_num == 0 || h.fps_den == 0) {
    f = {EBIV_ERR_CORRUPT, "ebiv: corrupt stream: frame rate " + std::to_string(h.fps_num) + "/" +
                               std::to_string(h.fps_den)};
    return false;
  }
  h.chroma_format = b[28];
  h.bit_depth = b[29];
  h.color_range = b[30];
  h.color_primaries = b[31];
  if (h.chroma_format != 0) {
    f = {EBIV_ERR_UNSUPPORTED, "ebiv: unsupported stream: chroma format " + std::to_string(h.chroma_format) +
                                   ", only 4:2:0 is supported"};
    return false;
  }
  if (h.bit_depth != 0) {
    f = {EBIV_ERR_UNSUPPORTED, "ebiv: unsupported stream: bit depth code " + std::to_string(h.bit_depth) +
                                   ", only 8-bit is supported"};
    return false;
  }
  h.tile_cols = le16(b + 32);
  h.tile_rows = le16(b + 34);
  h.flags = le32(b + 36);
  return true;
}

// validateEntry (reader.go:113-121); uint64 sums wrap as in Go.
bool valid_entry(const ebiv_index_entry& e, uint64_t payload_end) {
  if (e.offset < kFileHeaderSize || e.size == 0 || e.size > kMaxFrameSize) return false;
  return e.offset + kFrameRecordSize + static_cast<uint64_t>(e.size) <= payload_end;
}

// readIndex (reader.go:63-111): the footer's index, or false (the cue to scan).
bool read_index(const std::vector<uint8_t>& d, std::vector<ebiv_index_entry>& index) {
  const int64_t size = static_cast<int64_t>(d.size());
  if (size < static_cast<int64_t>(kFileHeaderSize + kFooterSize)) return false;
  const uint8_t* footer = d.data() + d.size() - kFooterSize;
  if (std::memcmp(footer + 16, "EBIX", 4) != 0) return false;
  const uint64_t offset = le64(footer);
  const uint32_t count = le32(footer + 8), want = le32(footer + 12);
  const int64_t end = static_cast<int64_t>(offset) + static_cast<int64_t>(count) * static_cast<int64_t>(kIndexEntrySize);
  if (offset < kFileHeaderSize || end != size - static_cast<int64_t>(kFooterSize)) return false;
  if (static_cast<int64_t>(offset) < 0) return false;  // Go's Seek to a negative position fails
  const uint8_t* buf = d.data() + offset;  // offset + count * 16 == size - 20 here
  if (crc32_ieee(buf, static_cast<size_t>(count) * kIndexEntrySize) != want) return false;
  index.resize(count);
  for (uint32_t i = 0; i < count; i++) {
    const uint8_t* p = buf + static_cast<size_t>(i) * kIndexEntrySize;
    index[i].offset = le64(p);
    index[i].size = le32(p + 8);
    index[i].keyframe = (le32(p + 12) & kFlagKeyframe) ? 1 : 0;
    if (!valid_entry(index[i], offset)) return false;
  }
  return !(count > 0 && !index[0].keyframe);
}

// scanFrames (reader.go:123-160): the records walked from the header on; a
// trailing partial record ends the scan without failing.
bool scan_frames(const std::vector<uint8_t>& d, std::vector<ebiv_index_entry>& index, Fail& f) {
  const int64_t size = static_cast<int64_t>(d.size());
  index.clear();
  int64_t off = kFileHeaderSize;
  while (off + static_cast<int64_t>(kFrameRecordSize) <= size) {
    const uint8_t* rec = d.data() + off;
    if (le16(rec) != kFrameSync) break;
    const bool key = (le16(rec + 2) & kFlagKeyframe) != 0;
    const uint32_t payload = le32(rec + 4);
    if (payload == 0 || payload > kMaxFrameSize) break;
    const int64_t end = off + static_cast<int64_t>(kFrameRecordSize) + payload;
    if (end > size) break;
    index.push_back(ebiv_index_entry{static_cast<uint64_t>(off), payload, key ? 1 : 0});
    off = end;
  }
  if (index.empty()) {
    f = {EBIV_ERR_CORRUPT, "ebiv: corrupt stream: no frames found"};
    return false;
  }
  if (!index[0].keyframe) {
    f = {EBIV_ERR_CORRUPT, "ebiv: corrupt stream: first frame is not a key frame"};
    return false;
  }
  return true;
}

void put_msg(char* msg, size_t cap, const std::string& s) {
  if (!msg || cap == 0) return;
  std::snprintf(msg, cap, "%s", s.c_str());
}

}  // namespace

struct ebiv_reader {
  std::vector<uint8_t> data;
  ebiv_file_header hdr{};
  std::vector<ebiv_index_entry> index;
  int64_t next = 0;
  bool closed = false;
  bool scanned = false;
};

extern "C" {

int ebiv_reader_open(const uint8_t* data, size_t len, ebiv_reader** out, char* msg, size_t msg_cap) {
  if (!out || (!data && len)) return EBIV_ERR_ARGUMENT;
  *out = nullptr;
  // io.ReadFull of the header: a short file is "not an EBIV file" (reader.go:39-44)
  if (len < kFileHeaderSize) {
    put_msg(msg, msg_cap, "ebiv: not an EBIV file");
    return EBIV_ERR_BAD_MAGIC;
  }
  std::unique_ptr<ebiv_reader> r(new (std::nothrow) ebiv_reader);
  if (!r) return EBIV_ERR_ARGUMENT;
  r->data.assign(data, data + len);
  Fail f{};
  if (!parse_header(r->data.data(), r->data.size(), r->hdr, f)) {
    put_msg(msg, msg_cap, f.msg);
    return f.code;
  }
  if (!read_index(r->data, r->index)) {
    if (!scan_frames(r->data, r->index, f)) {
      put_msg(msg, msg_cap, f.msg);
      return f.code;
    }
    r->scanned = true;
  }
  put_msg(msg, msg_cap, "");
  *out = r.release();
  return EBIV_OK;
}

void ebiv_reader_destroy(ebiv_reader* r) { delete r; }

int ebiv_reader_header(const ebiv_reader* r, ebiv_file_header* out) {
  if (!r || !out) return EBIV_ERR_ARGUMENT;
  *out = r->hdr;
  return EBIV_OK;
}

int64_t ebiv_reader_frame_count(const ebiv_reader* r) { return r ? static_cast<int64_t>(r->index.size()) : -1; }

int ebiv_reader_index(const ebiv_reader* r, ebiv_index_entry* out, size_t max, size_t* n) {
  if (!r || !n || (!out && max)) return EBIV_ERR_ARGUMENT;
  const size_t k = r->index.size() < max ? r->index.size() : max;
  if (k) std::memcpy(out, r->index.data(), k * sizeof(ebiv_index_entry));
  *n = r->index.size();
  return EBIV_OK;
}

int ebiv_reader_index_scanned(const ebiv_reader* r) { return r && r->scanned ? 1 : 0; }

int64_t ebiv_reader_duration(const ebiv_reader* r) {
  return r ? frame_time(r->hdr.fps_num, r->hdr.fps_den, static_cast<int64_t>(r->index.size())) : 0;
}

int ebiv_reader_next_packet(ebiv_reader* r, const uint8_t** data, size_t* len, int64_t* ts_ns, int32_t* keyframe) {
  if (!r || !data || !len) return EBIV_ERR_ARGUMENT;
  if (r->closed || r->next >= static_cast<int64_t>(r->index.size())) return EBIV_EOF;
  const ebiv_index_entry& e = r->index[static_cast<size_t>(r->next)];
  // The payload must lie inside the file (Go's read would fail otherwise).
  if (e.offset > r->data.size() || r->data.size() - e.offset < kFrameRecordSize + e.size) return EBIV_ERR_CORRUPT;
  *data = r->data.data() + e.offset + kFrameRecordSize;
  *len = e.size;
  if (ts_ns) *ts_ns = frame_time(r->hdr.fps_num, r->hdr.fps_den, r->next);
  if (keyframe) *keyframe = e.keyframe;
  r->next++;
  return EBIV_OK;
}

int ebiv_reader_seek(ebiv_reader* r, int64_t t_ns, int64_t* actual_ns) {
  if (!r) return EBIV_ERR_ARGUMENT;
  if (r->closed || r->index.empty()) return EBIV_EOF;
  int64_t target = frame_index_at(r->hdr.fps_num, r->hdr.fps_den, t_ns);
  const int64_t n = static_cast<int64_t>(r->index.size());
  if (target >= n) target = n - 1;
  while (target > 0 && !r->index[static_cast<size_t>(target)].keyframe) target--;  // keyframeBefore
  r->next = target;
  if (actual_ns) *actual_ns = frame_time(r->hdr.fps_num, r->hdr.fps_den, target);
  return EBIV_OK;
}

int ebiv_reader_close(ebiv_reader* r) {
  if (!r) return EBIV_ERR_ARGUMENT;
  r->closed = true;
  return EBIV_OK;
}

int64_t ebiv_frame_time(uint32_t fps_num, uint32_t fps_den, int64_t i) {
  return fps_num ? frame_time(fps_num, fps_den, i) : 0;
}

int64_t ebiv_frame_index_at(uint32_t fps_num, uint32_t fps_den, int64_t t_ns) {
  return fps_num && fps_den ? frame_index_at(fps_num, fps_den, t_ns) : 0;
}

}  // extern "C"
