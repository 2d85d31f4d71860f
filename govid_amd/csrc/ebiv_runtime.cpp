// ebiv_runtime.cpp — host runtime of the MI355X EBIV decoder behind the C ABI
// of include/ebiv_gpu.h.
//
// The host keeps what the reference keeps on the host side of Codec.Decode:
// the frame header (ebiv/format.go:376-409), the golden-override split
// (ebiv/codec.go:162-173), the coded header, table and tile framing
// (ebiv/frame.go:197-274, ebiv/rans.go:396-481) and the codec state machine
// (ebiv/codec.go:70-153).  Everything that touches pixels or the rANS streams
// runs on the device (ebiv_kernels.hip).
//
// Frame buffers live in a device pool.  A decode writes a free buffer; on
// success that buffer becomes both the ring slot's frame and the reference
// for the next inter frame, so the per-frame visible copy (frame.go:67-77)
// and the per-GOP golden copy (frame.go:42-46) of the reference are replaced
// by pointer moves.  Go's observable state is kept exactly: ring slots are
// consumed before decoding, failed decodes leave ref and the ring untouched,
// and a failed golden override leaves golden partially rewritten.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ebiv_gpu.h"
#include "ebiv_device.h"
#include "ebiv_format.h"

namespace ebiv {

hipError_t upload_tables(const int32_t dct[3][256], const uint8_t zz[3][256]);
void launch_parse(hipStream_t s, int ntiles, int max_tile_w, const uint8_t* pkt, const uint8_t* desc, MBRec* recs,
                  uint32_t* coefs, TileStatus* status);
void launch_recon_inter(hipStream_t s, int nmb, const uint8_t* desc, const MBRec* recs, const uint32_t* coefs,
                        const TileStatus* status, Planes rec, Planes ref, Planes gold);
void launch_recon_intra(hipStream_t s, int ntiles, const uint8_t* desc, const MBRec* recs, const uint32_t* coefs,
                        const TileStatus* status, Planes rec);
void launch_rgba(hipStream_t s, const uint8_t* y, int ys, const uint8_t* cb, const uint8_t* cr, int cs, int w, int h,
                 uint32_t* out);

namespace {

// ---- static tables (transform.go:27-40, scan.go:17-35, quant.go:19-29) ------
struct StaticTables {
  int32_t dct[3][256];
  uint8_t zz[3][256];
  int32_t qstep[kMaxQP + 1];
  StaticTables() {
    std::memset(dct, 0, sizeof dct);
    std::memset(zz, 0, sizeof zz);
    for (int s = 0; s < 3; s++) {
      const int n = 4 << s;
      int k = 0;
      for (int d = 0; d < 2 * n - 1; d++) {
        if ((d & 1) == 0) {
          for (int y = std::min(d, n - 1); y >= 0 && d - y < n; y--) zz[s][k++] = static_cast<uint8_t>(y * n + d - y);
        } else {
          for (int x = std::min(d, n - 1); x >= 0 && d - x < n; x--) zz[s][k++] = static_cast<uint8_t>((d - x) * n + x);
        }
      }
      for (int kk = 0; kk < n; kk++) {
        const double alpha = kk == 0 ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
        for (int i = 0; i < n; i++)
          dct[s][kk * n + i] = static_cast<int32_t>(std::lround(alpha * std::cos((2.0 * i + 1.0) * kk * M_PI / (2.0 * n)) * 8192.0));
      }
    }
    for (int qp = 0; qp <= kMaxQP; qp++) qstep[qp] = std::max<int32_t>(1, static_cast<int32_t>(std::lround(std::pow(2.0, qp / 6.0))));
  }
};
const StaticTables& statics() {
  static StaticTables t;
  return t;
}

std::mutex g_init_mu;
std::vector<bool> g_dev_ready;

int ensure_device(int dev, std::string& msg) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || dev < 0 || dev >= count) {
    msg = "no usable HIP device";
    return EBIV_ERR_DEVICE;
  }
  if (static_cast<int>(g_dev_ready.size()) < count) g_dev_ready.resize(count, false);
  if (hipSetDevice(dev) != hipSuccess) { msg = "hipSetDevice failed"; return EBIV_ERR_DEVICE; }
  if (!g_dev_ready[dev]) {
    const StaticTables& t = statics();
    if (upload_tables(t.dct, t.zz) != hipSuccess) { msg = "constant table upload failed"; return EBIV_ERR_DEVICE; }
    g_dev_ready[dev] = true;
  }
  return EBIV_OK;
}

// prevFreqs slot: Go keeps the backing array across key-frame wipes
// (rans.go:397-405, 473-481), so a re-grown slice shows its old values.
struct PrevFreq {
  uint32_t v[kMaxAlphabet];
  int len = 0;
  bool alloc = false;
};

struct DevBuf {
  uint8_t* base = nullptr;
  Planes p{};
};

template <typename T>
struct DevArray {
  T* d = nullptr;
  size_t cap = 0;
  bool reserve(size_t n) {
    if (n <= cap) return true;
    if (d) hipFree(d);
    d = nullptr;
    cap = 0;
    if (hipMalloc(&d, n * sizeof(T)) != hipSuccess) { d = nullptr; return false; }
    cap = n;
    return true;
  }
  void release() { if (d) hipFree(d); d = nullptr; cap = 0; }
};

template <typename T>
struct PinnedArray {
  T* h = nullptr;
  size_t cap = 0;
  bool reserve(size_t n) {
    if (n <= cap) return true;
    if (h) hipHostFree(h);
    h = nullptr;
    cap = 0;
    if (hipHostMalloc(&h, n * sizeof(T), hipHostMallocDefault) != hipSuccess) { h = nullptr; return false; }
    cap = n;
    return true;
  }
  void release() { if (h) hipHostFree(h); h = nullptr; cap = 0; }
};

// One coded payload ready for the device: descriptor + launch geometry.
struct CodedJob {
  int ntiles = 0;
  int max_tile_w = 0;
  size_t desc_bytes = 0;
};

}  // namespace
}  // namespace ebiv

using namespace ebiv;

struct ebiv_codec {
  int ring_size = 3;
  int device = 0;
  bool ready = false;
  int W = 0, H = 0, CW = 0, CH = 0, mbc = 0, mbr = 0;
  int ystride = 0, cstride = 0;  // host ring image strides (geometry.go:22-30)
  int next = 0;
  std::vector<DevBuf> pool;
  std::vector<int> ring;  // pool index per slot, -1 = never written
  int ref = -1, golden = -1, work = -1;
  std::vector<uint8_t*> host_ring;
  PrevFreq prev[kNumContexts];
  hipStream_t stream = nullptr;
  DevArray<uint8_t> d_pkt;
  PinnedArray<uint8_t> h_pkt;
  DevArray<uint8_t> d_desc[2];
  PinnedArray<uint8_t> h_desc[2];
  DevArray<MBRec> d_recs[2];
  DevArray<uint32_t> d_coefs[2];
  DevArray<TileStatus> d_status[2];
  PinnedArray<TileStatus> h_status[2];
  DevArray<uint32_t> d_rgba;
  std::string msg;
  ebiv_stats stats{0, 0, 0};
  // Kernel timing: events bracket each launch while enabled.
  bool timing = false;
  hipEvent_t ev[2][4] = {};
  int ev_used[2] = {0, 0};
  ebiv_kernel_times ktimes{};

  int fail(int code, const std::string& m) {
    msg = m;
    return code;
  }
  ~ebiv_codec() {
    if (stream) hipStreamSynchronize(stream);
    for (auto& b : pool) if (b.base) hipFree(b.base);
    for (auto* h : host_ring) if (h) hipHostFree(h);
    d_pkt.release();
    h_pkt.release();
    for (int i = 0; i < 2; i++) {
      d_desc[i].release(); h_desc[i].release(); d_recs[i].release(); d_coefs[i].release();
      d_status[i].release(); h_status[i].release();
    }
    d_rgba.release();
    for (auto& row : ev)
      for (auto& e : row)
        if (e) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
  }

  bool ensure_events() {
    for (auto& row : ev)
      for (auto& e : row)
        if (!e && hipEventCreate(&e) != hipSuccess) return false;
    return true;
  }
  // Adds the timed launches of slot `slot` (after the stream has synchronised).
  void collect_times(int slot, bool inter) {
    if (!timing || !ev_used[slot]) return;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev[slot][0], ev[slot][1]) == hipSuccess) { ktimes.parse_ms += ms; ktimes.parse_n++; }
    if (inter && hipEventElapsedTime(&ms, ev[slot][1], ev[slot][2]) == hipSuccess) { ktimes.inter_ms += ms; ktimes.inter_n++; }
    if (hipEventElapsedTime(&ms, ev[slot][2], ev[slot][3]) == hipSuccess) { ktimes.intra_ms += ms; ktimes.intra_n++; }
    ev_used[slot] = 0;
  }

  // configure, codec.go:129-153.
  int configure(int w, int h) {
    if (static_cast<long>(w) * h > kMaxDecodePixels)
      return fail(EBIV_ERR_UNSUPPORTED, "ebiv: unsupported stream feature: frame exceeds the decoder's pixel bound");
    if (ready) {
      if (w != W || h != H) return fail(EBIV_ERR_DIMENSIONS, "ebiv: frame dimensions do not match the stream");
      return EBIV_OK;
    }
    W = w; H = h; CW = (w + 1) / 2; CH = (h + 1) / 2;
    mbc = (w + kMB - 1) / kMB;
    mbr = (h + kMB - 1) / kMB;
    ystride = align_up(W, 64);
    cstride = align_up(CW, 64);
    const int pw = mbc * kMB, ph = mbr * kMB, cw = mbc * kCMB, ch = mbr * kCMB;
    const size_t ysz = static_cast<size_t>(pw) * ph, csz = static_cast<size_t>(align_up(cw * ch, 256));
    pool.assign(static_cast<size_t>(ring_size) + 4, DevBuf{});
    for (auto& b : pool) {
      if (hipMalloc(&b.base, ysz + 2 * csz) != hipSuccess) return fail(EBIV_ERR_DEVICE, "frame pool allocation failed");
      if (hipMemsetAsync(b.base, 0, ysz + 2 * csz, stream) != hipSuccess) return fail(EBIV_ERR_DEVICE, "frame pool clear failed");
      b.p = Planes{b.base, b.base + ysz, b.base + ysz + csz, pw, ph, cw, ch};
    }
    ring.assign(ring_size, -1);
    ref = 0; golden = 1; work = 2;
    host_ring.assign(ring_size, nullptr);
    const size_t hsz = static_cast<size_t>(ystride) * H + 2 * static_cast<size_t>(cstride) * CH;
    for (auto*& hr : host_ring) {
      if (hipHostMalloc(reinterpret_cast<void**>(&hr), hsz, hipHostMallocDefault) != hipSuccess)
        return fail(EBIV_ERR_DEVICE, "host ring allocation failed");
      std::memset(hr, 0, hsz);
    }
    const size_t nmb = static_cast<size_t>(mbc) * mbr;
    for (int i = 0; i < 2; i++) {
      if (!d_recs[i].reserve(nmb) || !d_coefs[i].reserve(nmb * kCoefPerMB))
        return fail(EBIV_ERR_DEVICE, "macroblock buffer allocation failed");
    }
    if (hipStreamSynchronize(stream) != hipSuccess) return fail(EBIV_ERR_DEVICE, "frame pool clear failed");
    next = 0;
    ready = true;
    return EBIV_OK;
  }

  bool referenced(int b, int extra = -1) const {
    if (b == ref || b == golden || b == work || b == extra) return true;
    for (int r : ring) if (r == b) return true;
    return false;
  }
  int free_buffer(int extra = -1) const {
    for (int i = 0; i < static_cast<int>(pool.size()); i++) if (!referenced(i, extra)) return i;
    return -1;
  }

  // parseTablesInto (rans.go:396-471) into the frame header's cumulative
  // tables, updating prev exactly as Go does (including on failure).
  int parse_tables(const uint8_t* b, size_t len, bool key, FrameHdr& hdr, size_t* consumed) {
    if (key) for (auto& p : prev) p.len = 0;
    size_t pos = 0;
    hdr.used[0] = hdr.used[1] = hdr.used[2] = 0;
    auto install = [&](int c, const uint32_t* f, int n) -> bool {
      uint32_t sum = 0;
      for (int s = 0; s < n; s++) sum += f[s];
      if (sum != kRansM) return false;
      // Go indexes slot2sym with start+k; an entry above ransM whose wrapped
      // sum still reaches ransM would panic there.  Such a stream is corrupt.
      for (int s = 0; s < n; s++) if (f[s] > kRansM) return false;
      uint32_t cum = 0;
      for (int s = 0; s < kCumStride; s++) {
        if (s < n) { hdr.cum[c][s] = static_cast<uint16_t>(cum); cum += f[s]; }
        else if (s == n) hdr.cum[c][s] = static_cast<uint16_t>(kRansM);
        else hdr.cum[c][s] = 0xFFFF;
      }
      hdr.used[c >> 5] |= 1u << (c & 31);
      return true;
    };
    for (int c = 0; c < kNumContexts; c++) {
      const int n = alphabet_size(c);
      if (pos >= len) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: truncated table modes");
      const uint8_t mode = b[pos++];
      if (c == kCtxBypass && mode != 0) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: bypass context must not ship a table");
      const bool have_ref = prev[c].len == n;
      PrevFreq& p = prev[c];
      switch (mode) {
        case 0:
          break;
        case 1:
          if (key || !have_ref) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: context references a previous table that does not exist");
          if (!install(c, p.v, n)) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: context frequencies do not sum to 4096");
          break;
        case 2: {
          if (!p.alloc) { std::memset(p.v, 0, sizeof p.v); p.alloc = true; }
          p.len = n;
          for (int s = 0; s < n; s++) {
            uint64_t v = 0;
            const int k = uvarint(b + pos, len - pos, &v);
            if (k <= 0) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: truncated frequency table body");
            pos += static_cast<size_t>(k);
            p.v[s] = static_cast<uint32_t>(v);
          }
          if (!install(c, p.v, n)) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: context frequencies do not sum to 4096");
          break;
        }
        case 3:
          if (key || !have_ref) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: context deltas a previous table that does not exist");
          for (int s = 0; s < n; s++) {
            int64_t v = 0;
            const int k = varint(b + pos, len - pos, &v);
            if (k <= 0) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: truncated frequency table delta");
            pos += static_cast<size_t>(k);
            const int64_t nv = static_cast<int64_t>(p.v[s]) + v;
            if (nv < 0 || nv > static_cast<int64_t>(kRansM)) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: context delta out of range");
            p.v[s] = static_cast<uint32_t>(nv);
          }
          if (!install(c, p.v, n)) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: context frequencies do not sum to 4096");
          break;
        default:
          return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: unknown table mode");
      }
    }
    const uint32_t bypass[2] = {kRansM / 2, kRansM / 2};  // installBypass
    install(kCtxBypass, bypass, 2);
    *consumed = pos;
    return EBIV_OK;
  }

  // decodeState.decodeCoded up to the tile fan-out (decode_intra.go:55-64,
  // frame.go:197-274): fills the host descriptor of slot `slot`.  body is a
  // sub-range of the packet starting at pkt_base.
  int prepare_coded(const uint8_t* pkt_base, const uint8_t* body, size_t len, bool inter, bool key, int slot,
                    CodedJob& job) {
    if (len < 1) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: coded header truncated");
    const int qp = body[0];
    size_t pos = 1;
    uint64_t cols = 0, rows = 0, tlen = 0;
    int n = uvarint(body + pos, len - pos, &cols);
    if (n <= 0) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: coded header tile columns");
    pos += static_cast<size_t>(n);
    n = uvarint(body + pos, len - pos, &rows);
    if (n <= 0) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: coded header tile rows");
    pos += static_cast<size_t>(n);
    if (qp > kMaxQP || cols == 0 || rows == 0 || cols > 65535 || rows > 65535)
      return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: coded header values out of range");
    n = uvarint(body + pos, len - pos, &tlen);
    if (n <= 0 || tlen > static_cast<uint64_t>(len - pos - static_cast<size_t>(n)))
      return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: bad table length");
    pos += static_cast<size_t>(n);

    const int gc = static_cast<int>(std::min<uint64_t>(cols, static_cast<uint64_t>(mbc)));
    const int gr = static_cast<int>(std::min<uint64_t>(rows, static_cast<uint64_t>(mbr)));
    const int ntiles_grid = gc * gr;
    const size_t dsize = desc_size(ntiles_grid, mbc, mbr);
    if (!h_desc[slot].reserve(dsize) || !d_desc[slot].reserve(dsize)) return fail(EBIV_ERR_DEVICE, "descriptor allocation failed");
    uint8_t* hd = h_desc[slot].h;
    std::memset(hd, 0, dsize);
    FrameHdr& hdr = *reinterpret_cast<FrameHdr*>(hd);
    size_t consumed = 0;
    int rc = parse_tables(body + pos, static_cast<size_t>(tlen), key, hdr, &consumed);
    if (rc) return rc;
    if (consumed != tlen) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: trailing table bytes");
    pos += static_cast<size_t>(tlen);

    const uint64_t want = cols * rows;
    std::vector<std::pair<size_t, size_t>> tiles;
    while (tiles.size() < want) {
      uint64_t l = 0;
      n = uvarint(body + pos, len - pos, &l);
      if (n <= 0 || l > static_cast<uint64_t>(len - pos - static_cast<size_t>(n)))
        return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: bad tile length");
      pos += static_cast<size_t>(n);
      tiles.emplace_back(static_cast<size_t>(body + pos - pkt_base), static_cast<size_t>(l));
      pos += static_cast<size_t>(l);
    }
    if (static_cast<uint64_t>(ntiles_grid) != tiles.size()) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream");

    hdr.mbc = mbc;
    hdr.mbr = mbr;
    hdr.ntiles = ntiles_grid;
    hdr.inter = inter ? 1 : 0;
    const int q = qp;
    hdr.qac = statics().qstep[q];
    hdr.qdc = q >= 6 ? statics().qstep[q - 6] : statics().qstep[q];
    hdr.grid_cols = gc;
    hdr.grid_rows = gr;
    TileDesc* td = reinterpret_cast<TileDesc*>(hd + desc_tiles_off());
    uint16_t* ct = reinterpret_cast<uint16_t*>(hd + desc_coltile_off(ntiles_grid));
    uint16_t* rt = reinterpret_cast<uint16_t*>(hd + desc_rowtile_off(ntiles_grid, mbc));
    uint32_t coef = 0;
    int maxw = 1;
    for (int i = 0; i < ntiles_grid; i++) {  // tileGrid.bounds, frame.go:170-177
      const int tc = i % gc, tr = i / gc;
      TileDesc& t = td[i];
      t.byte_off = static_cast<uint32_t>(tiles[i].first);
      t.byte_len = static_cast<uint32_t>(tiles[i].second);
      t.x0 = static_cast<uint16_t>(mbc * tc / gc);
      t.x1 = static_cast<uint16_t>(mbc * (tc + 1) / gc);
      t.y0 = static_cast<uint16_t>(mbr * tr / gr);
      t.y1 = static_cast<uint16_t>(mbr * (tr + 1) / gr);
      t.coef_base = coef;
      t.pad = 0;
      coef += static_cast<uint32_t>((t.x1 - t.x0) * (t.y1 - t.y0) * kCoefPerMB);
      maxw = std::max(maxw, t.x1 - t.x0);
    }
    for (int tc = 0; tc < gc; tc++)
      for (int x = mbc * tc / gc; x < mbc * (tc + 1) / gc; x++) ct[x] = static_cast<uint16_t>(tc);
    for (int tr = 0; tr < gr; tr++)
      for (int y = mbr * tr / gr; y < mbr * (tr + 1) / gr; y++) rt[y] = static_cast<uint16_t>(tr);
    hdr.max_tile_w = maxw;
    job.ntiles = ntiles_grid;
    job.max_tile_w = maxw;
    job.desc_bytes = dsize;
    if (!d_status[slot].reserve(ntiles_grid) || !h_status[slot].reserve(ntiles_grid))
      return fail(EBIV_ERR_DEVICE, "status allocation failed");
    return EBIV_OK;
  }

  // Queues the device decode of a prepared payload into `rec`.
  int launch(int slot, const CodedJob& job, const uint8_t* dpkt, bool inter, const Planes& rec, const Planes& r,
             const Planes& g) {
    if (hipMemcpyAsync(d_desc[slot].d, h_desc[slot].h, job.desc_bytes, hipMemcpyHostToDevice, stream) != hipSuccess)
      return fail(EBIV_ERR_DEVICE, "descriptor upload failed");
    const bool tm = timing && ensure_events();
    if (tm) hipEventRecord(ev[slot][0], stream);
    launch_parse(stream, job.ntiles, job.max_tile_w, dpkt, d_desc[slot].d, d_recs[slot].d, d_coefs[slot].d, d_status[slot].d);
    if (tm) hipEventRecord(ev[slot][1], stream);
    if (inter)
      launch_recon_inter(stream, mbc * mbr, d_desc[slot].d, d_recs[slot].d, d_coefs[slot].d, d_status[slot].d, rec, r, g);
    if (tm) hipEventRecord(ev[slot][2], stream);
    launch_recon_intra(stream, job.ntiles, d_desc[slot].d, d_recs[slot].d, d_coefs[slot].d, d_status[slot].d, rec);
    if (tm) {
      hipEventRecord(ev[slot][3], stream);
      ev_used[slot] = 1;
    }
    if (hipMemcpyAsync(h_status[slot].h, d_status[slot].d, sizeof(TileStatus) * job.ntiles, hipMemcpyDeviceToHost,
                       stream) != hipSuccess)
      return fail(EBIV_ERR_DEVICE, "status download failed");
    if (hipGetLastError() != hipSuccess) return fail(EBIV_ERR_DEVICE, "kernel launch failed");
    return EBIV_OK;
  }

  // First failing tile in tile order (decode_intra.go:86-93).
  int tile_result(int slot, const CodedJob& job) {
    for (int i = 0; i < job.ntiles; i++) {
      const TileStatus& s = h_status[slot].h[i];
      stats.symbols += s.nsym;
      stats.tiles++;
    }
    for (int i = 0; i < job.ntiles; i++) {
      const TileStatus& s = h_status[slot].h[i];
      if (s.err) {
        switch (s.err >> 8) {
          case kTileShort: return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: tile stream too short to initialize rANS");
          case kTileUnused: return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: symbol decoded from unused context");
          default: return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: rANS stream underrun");
        }
      }
    }
    return EBIV_OK;
  }

  void fill_frame(int b, ebiv_frame* out, int64_t ts, bool on_device, int slot) {
    const DevBuf& f = pool[b];
    out->width = W; out->height = H; out->cwidth = CW; out->cheight = CH;
    out->timestamp_ns = ts;
    out->ring_slot = slot;
    if (on_device) {
      out->y = f.p.y; out->cb = f.p.cb; out->cr = f.p.cr;
      out->y_stride = f.p.pw; out->c_stride = f.p.cw;
      out->on_device = 1;
    } else {
      uint8_t* h = host_ring[slot];
      out->y = h;
      out->cb = h + static_cast<size_t>(ystride) * H;
      out->cr = out->cb + static_cast<size_t>(cstride) * CH;
      out->y_stride = ystride; out->c_stride = cstride;
      out->on_device = 0;
    }
  }

  // storeImage into the host ring (frame.go:67-77): visible region only.
  int download(int b, int slot) {
    const DevBuf& f = pool[b];
    uint8_t* h = host_ring[slot];
    uint8_t* hcb = h + static_cast<size_t>(ystride) * H;
    uint8_t* hcr = hcb + static_cast<size_t>(cstride) * CH;
    if (hipMemcpy2DAsync(h, ystride, f.p.y, f.p.pw, W, H, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipMemcpy2DAsync(hcb, cstride, f.p.cb, f.p.cw, CW, CH, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipMemcpy2DAsync(hcr, cstride, f.p.cr, f.p.cw, CW, CH, hipMemcpyDeviceToHost, stream) != hipSuccess)
      return fail(EBIV_ERR_DEVICE, "frame download failed");
    if (hipStreamSynchronize(stream) != hipSuccess) return fail(EBIV_ERR_DEVICE, "frame download failed");
    return EBIV_OK;
  }

  // Codec.Decode, codec.go:70-111.
  int decode(const uint8_t* data, const uint8_t* dev_data, size_t len, int64_t ts, ebiv_frame* out) {
    if (!data && len) return fail(EBIV_ERR_ARGUMENT, "null packet");
    if (hipSetDevice(device) != hipSuccess) return fail(EBIV_ERR_DEVICE, "hipSetDevice failed");
    // parseFrameHeader, format.go:376-409
    if (len < 1) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: empty frame payload");
    const int type = data[0] & 3, coding = (data[0] >> 2) & 7;
    const bool gov = ((data[0] >> 5) & 1) != 0;
    if (type != 0 && type != 1) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: frame type");
    if (gov && type != 0) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: golden override on a non-key frame");
    size_t hlen = 1;
    if (type == 0) {
      if (len < 6) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: key frame header truncated");
      const int w = data[1] | data[2] << 8, h = data[3] | data[4] << 8;
      if (w == 0 || h == 0) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: key frame dimensions");
      if ((data[5] >> 4) != 0) return fail(EBIV_ERR_UNSUPPORTED, "ebiv: unsupported stream feature: chroma format");
      if ((data[5] & 15) != 0) return fail(EBIV_ERR_UNSUPPORTED, "ebiv: unsupported stream feature: bit depth code");
      hlen = 6;
      const int rc = configure(w, h);
      if (rc) return rc;
    } else if (!ready) {
      return fail(EBIV_ERR_NOKEYFRAME, "ebiv: inter frame before any key frame");
    }
    const int slot = next;
    if (++next == ring_size) next = 0;
    const bool on_device = dev_data != nullptr;

    // The packet on the device (the kernels read tile bytes from it).
    const uint8_t* dpkt = dev_data;
    if (!on_device && (coding == 1 || coding == 2 || coding == 0)) {
      if (!d_pkt.reserve(len + 256) || !h_pkt.reserve(len + 256)) return fail(EBIV_ERR_DEVICE, "packet buffer allocation failed");
      std::memcpy(h_pkt.h, data, len);
      if (hipMemcpyAsync(d_pkt.d, h_pkt.h, len, hipMemcpyHostToDevice, stream) != hipSuccess)
        return fail(EBIV_ERR_DEVICE, "packet upload failed");
      dpkt = d_pkt.d;
    }

    if (coding == 0) {  // decodeRaw, codec.go:195-207
      const size_t want = static_cast<size_t>(W) * H + 2 * static_cast<size_t>(CW) * CH;
      if (len - hlen != want) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: raw frame size");
      const int b = free_buffer();
      const DevBuf& f = pool[b];
      const uint8_t* src = dpkt + hlen;
      if (hipMemcpy2DAsync(f.p.y, f.p.pw, src, W, W, H, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
          hipMemcpy2DAsync(f.p.cb, f.p.cw, src + static_cast<size_t>(W) * H, CW, CW, CH, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
          hipMemcpy2DAsync(f.p.cr, f.p.cw, src + static_cast<size_t>(W) * H + static_cast<size_t>(CW) * CH, CW, CW, CH,
                           hipMemcpyDeviceToDevice, stream) != hipSuccess)
        return fail(EBIV_ERR_DEVICE, "raw frame copy failed");
      ring[slot] = b;
      if (!on_device) {
        int rc = download(b, slot);
        if (rc) return rc;
      } else if (hipStreamSynchronize(stream) != hipSuccess) {
        return fail(EBIV_ERR_DEVICE, "raw frame copy failed");
      }
      stats.frames++;
      fill_frame(b, out, ts, on_device, slot);
      return EBIV_OK;
    }
    if (coding != 1 && coding != 2) return fail(EBIV_ERR_UNSUPPORTED, "ebiv: unsupported stream feature: coding mode");

    // Codec.decodeCoded, codec.go:162-192.
    const bool inter = coding == 2, key = type == 0;
    const uint8_t* body = data + hlen;
    size_t blen = len - hlen;
    const uint8_t* ov = nullptr;
    size_t ovlen = 0;
    if (gov) {
      uint64_t kl = 0;
      const int k = uvarint(body, blen, &kl);
      if (k <= 0 || kl > static_cast<uint64_t>(blen - static_cast<size_t>(k)))
        return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: golden override split");
      ov = body + k + kl;
      ovlen = blen - static_cast<size_t>(k) - static_cast<size_t>(kl);
      body += k;
      blen = static_cast<size_t>(kl);
    }
    CodedJob job;
    int rc = prepare_coded(data, body, blen, inter, key, 0, job);
    if (rc) return rc;
    const Planes& wp = pool[work].p;
    rc = launch(0, job, dpkt, inter, wp, pool[ref].p, pool[golden].p);
    if (rc) return rc;

    // The override is decoded right behind the key frame; its table parse is
    // only kept if the key frame itself decodes (Go parses it afterwards).
    CodedJob ovjob;
    int gtarget = -1;
    PrevFreq saved[kNumContexts];
    bool ov_prepared = false;
    int ov_rc = EBIV_OK;
    std::string ov_msg;
    if (key && ov) {
      std::memcpy(saved, prev, sizeof prev);
      ov_rc = prepare_coded(data, ov, ovlen, true, false, 1, ovjob);
      ov_msg = msg;
      if (ov_rc == EBIV_OK) {
        // Decode into a fresh buffer holding a copy of the old golden: if the
        // key frame fails it is dropped (Go never reaches the override), and
        // if the override fails it is exactly Go's partially rewritten golden.
        gtarget = free_buffer();
        const DevBuf& src = pool[golden];
        const DevBuf& dst = pool[gtarget];
        const size_t bytes = static_cast<size_t>(src.p.pw) * src.p.ph + 2 * static_cast<size_t>(align_up(src.p.cw * src.p.ch, 256));
        if (hipMemcpyAsync(dst.base, src.base, bytes, hipMemcpyDeviceToDevice, stream) != hipSuccess)
          return fail(EBIV_ERR_DEVICE, "golden copy failed");
        rc = launch(1, ovjob, dpkt, true, pool[gtarget].p, wp, wp);
        if (rc) return rc;
        ov_prepared = true;
      }
    }
    if (hipStreamSynchronize(stream) != hipSuccess) return fail(EBIV_ERR_DEVICE, "decode failed on the device");
    collect_times(0, inter);
    collect_times(1, true);
    rc = tile_result(0, job);
    if (rc) {
      if (key && ov) std::memcpy(prev, saved, sizeof prev);
      return rc;
    }
    if (key && ov) {
      if (ov_rc) return fail(ov_rc, ov_msg);
      if (ov_prepared) {
        golden = gtarget;  // on failure golden keeps the partial rewrite
        rc = tile_result(1, ovjob);
        if (rc) return rc;
      }
    } else if (key) {
      golden = work;  // golden.copyFrom(work) by sharing the buffer
    }
    ring[slot] = work;
    ref = work;
    work = -1;
    work = free_buffer();
    stats.frames++;
    if (!on_device) {
      rc = download(ref, slot);
      if (rc) return rc;
    }
    fill_frame(ref, out, ts, on_device, slot);
    return EBIV_OK;
  }
};

extern "C" {

int ebiv_codec_create(int ring_size, int device, ebiv_codec** out) {
  if (!out) return EBIV_ERR_ARGUMENT;
  *out = nullptr;
  std::string msg;
  int rc = ensure_device(device, msg);
  if (rc) return rc;
  auto* c = new ebiv_codec();
  c->ring_size = ring_size >= 2 ? ring_size : 3;
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return EBIV_ERR_DEVICE;
  }
  *out = c;
  return EBIV_OK;
}

int ebiv_codec_decode(ebiv_codec* c, const uint8_t* data, size_t len, int64_t ts, ebiv_frame* out) {
  if (!c || !out) return EBIV_ERR_ARGUMENT;
  return c->decode(data, nullptr, len, ts, out);
}

int ebiv_codec_decode_device(ebiv_codec* c, const uint8_t* host_data, const uint8_t* dev_data, size_t len, int64_t ts,
                             ebiv_frame* out) {
  if (!c || !out || (!dev_data && len)) return EBIV_ERR_ARGUMENT;
  return c->decode(host_data, dev_data, len, ts, out);
}

void ebiv_codec_flush(ebiv_codec* c) {
  if (c) c->next = 0;
}

void ebiv_codec_destroy(ebiv_codec* c) { delete c; }

const char* ebiv_codec_last_error(const ebiv_codec* c) { return c ? c->msg.c_str() : "null codec"; }

int ebiv_codec_stats(const ebiv_codec* c, ebiv_stats* out) {
  if (!c || !out) return EBIV_ERR_ARGUMENT;
  *out = c->stats;
  return EBIV_OK;
}

int ebiv_codec_debug_dump(ebiv_codec* c, int slot, void* recs, size_t recs_bytes, uint32_t* coefs, size_t coef_count,
                          void* status, size_t status_bytes) {
  if (!c || slot < 0 || slot > 1 || !c->ready) return -EBIV_ERR_ARGUMENT;
  if (hipSetDevice(c->device) != hipSuccess) return -EBIV_ERR_DEVICE;
  const size_t nmb = static_cast<size_t>(c->mbc) * c->mbr;
  if (recs && hipMemcpy(recs, c->d_recs[slot].d, std::min(recs_bytes, nmb * sizeof(MBRec)), hipMemcpyDeviceToHost) != hipSuccess)
    return -EBIV_ERR_DEVICE;
  if (coefs && hipMemcpy(coefs, c->d_coefs[slot].d, std::min(coef_count, nmb * kCoefPerMB) * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return -EBIV_ERR_DEVICE;
  if (status && c->d_status[slot].d &&
      hipMemcpy(status, c->d_status[slot].d, std::min(status_bytes, c->d_status[slot].cap * sizeof(TileStatus)),
                hipMemcpyDeviceToHost) != hipSuccess)
    return -EBIV_ERR_DEVICE;
  return static_cast<int>(nmb);
}

int ebiv_codec_set_timing(ebiv_codec* c, int enabled) {
  if (!c) return EBIV_ERR_ARGUMENT;
  c->timing = enabled != 0;
  c->ktimes = ebiv_kernel_times{};
  return EBIV_OK;
}

int ebiv_codec_kernel_times(const ebiv_codec* c, ebiv_kernel_times* out) {
  if (!c || !out) return EBIV_ERR_ARGUMENT;
  *out = c->ktimes;
  return EBIV_OK;
}

int ebiv_frame_rgba(ebiv_codec* c, const ebiv_frame* f, uint8_t* dst, int dst_on_device) {
  if (!c || !f || !dst || !c->ready) return EBIV_ERR_ARGUMENT;
  if (f->ring_slot < 0 || f->ring_slot >= c->ring_size || c->ring[f->ring_slot] < 0) return EBIV_ERR_ARGUMENT;
  if (hipSetDevice(c->device) != hipSuccess) return EBIV_ERR_DEVICE;
  const DevBuf& b = c->pool[c->ring[f->ring_slot]];
  const size_t px = static_cast<size_t>(c->W) * c->H;
  uint32_t* out = reinterpret_cast<uint32_t*>(dst);
  if (!dst_on_device) {
    if (!c->d_rgba.reserve(px)) return EBIV_ERR_DEVICE;
    out = c->d_rgba.d;
  }
  const bool tm = c->timing && c->ensure_events();
  if (tm) hipEventRecord(c->ev[0][0], c->stream);
  launch_rgba(c->stream, b.p.y, b.p.pw, b.p.cb, b.p.cr, b.p.cw, c->W, c->H, out);
  if (tm) hipEventRecord(c->ev[0][1], c->stream);
  if (hipGetLastError() != hipSuccess) return EBIV_ERR_DEVICE;
  if (!dst_on_device &&
      hipMemcpyAsync(dst, c->d_rgba.d, px * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    return EBIV_ERR_DEVICE;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return EBIV_ERR_DEVICE;
  if (tm) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev[0][0], c->ev[0][1]) == hipSuccess) { c->ktimes.rgba_ms += ms; c->ktimes.rgba_n++; }
  }
  return EBIV_OK;
}

int ebiv_device_alloc(int device, size_t bytes, void** out) {
  if (!out) return EBIV_ERR_ARGUMENT;
  std::string msg;
  int rc = ensure_device(device, msg);
  if (rc) return rc;
  return hipMalloc(out, bytes + 256) == hipSuccess ? EBIV_OK : EBIV_ERR_DEVICE;
}
int ebiv_device_free(void* p) { return hipFree(p) == hipSuccess ? EBIV_OK : EBIV_ERR_DEVICE; }
int ebiv_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? EBIV_OK : EBIV_ERR_DEVICE;
}
int ebiv_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? EBIV_OK : EBIV_ERR_DEVICE;
}
int ebiv_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

}  // extern "C"
