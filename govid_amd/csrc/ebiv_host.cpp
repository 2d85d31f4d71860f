// ebiv_host.cpp — the host half of Codec.Decode (see ebiv_host.h).
#include "ebiv_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>

namespace ebiv {

namespace {

// quantStep (quant.go:19-29): max(1, round(2^(qp/6))).  Every value is at
// least 0.0087 away from a rounding tie, so any correct libm gives Go's.
struct QSteps {
  int32_t v[kMaxQP + 1];
  QSteps() {
    for (int qp = 0; qp <= kMaxQP; qp++) v[qp] = std::max<int32_t>(1, static_cast<int32_t>(std::lround(std::pow(2.0, qp / 6.0))));
  }
};
const QSteps& qsteps() {
  static QSteps q;
  return q;
}

}  // namespace

int HostCodec::configure(int w, int h) {  // codec.go:129-153
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
  if (int rc = allocate_frames()) return rc;  // (ready stays false: the next key tries again)
  next = 0;
  ready = true;
  return EBIV_OK;
}

int HostCodec::parse_tables(const uint8_t* b, size_t len, bool key, FrameHdr& hdr, size_t* consumed) {
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
    int vec = 0, lane0 = 0;
    if (table_place(c, &vec, &lane0)) {
      uint32_t start = 0;
      for (int s = 0; s < n; s++) {
        hdr.vt[vec][lane0 + s] = start << 16 | f[s];
        start += f[s];
      }
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

int HostCodec::prepare_coded(const uint8_t* pkt_base, const uint8_t* body, size_t len, bool inter, bool key, int si) {
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
  const int nt = gc * gr;
  if (static_cast<int>(payloads.size()) <= si) payloads.resize(static_cast<size_t>(si) + 1);
  HostPayload& hp = payloads[static_cast<size_t>(si)];
  // A context no table installs reads as kNoSym on every lane of the
  // entropy kernel's table search, which finds no symbol there: the
  // "symbol decoded from unused context" error of rans.go:275-278.
  FrameHdr hdr;
  std::memset(&hdr, 0, sizeof hdr);
  std::memset(hdr.vt, 0xFF, sizeof hdr.vt);
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
  if (static_cast<uint64_t>(nt) != tiles.size()) return fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream");

  hdr.mbc = mbc;
  hdr.mbr = mbr;
  hdr.ntiles = nt;
  hdr.inter = inter ? 1 : 0;
  hdr.qac = qsteps().v[qp];
  hdr.qdc = qp >= 6 ? qsteps().v[qp - 6] : qsteps().v[qp];
  hdr.grid_cols = gc;
  hdr.grid_rows = gr;
  hp.desc.assign(desc_size(nt, mbc, mbr), 0);
  uint8_t* hd = hp.desc.data();
  TileDesc* td = reinterpret_cast<TileDesc*>(hd + desc_tiles_off());
  uint16_t* ct = reinterpret_cast<uint16_t*>(hd + desc_coltile_off(nt));
  uint16_t* rt = reinterpret_cast<uint16_t*>(hd + desc_rowtile_off(nt, mbc));
  uint32_t coef = 0;
  int maxw = 1;
  for (int i = 0; i < nt; i++) {  // tileGrid.bounds, frame.go:170-177
    const int tc = i % gc, tr = i / gc;
    TileDesc& t = td[i];
    t.byte_off = static_cast<uint32_t>(tiles[static_cast<size_t>(i)].first);
    t.byte_len = static_cast<uint32_t>(tiles[static_cast<size_t>(i)].second);
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
  std::memcpy(hd, &hdr, sizeof hdr);
  hp.ntiles = nt;
  hp.max_tile_w = maxw;
  hp.inter = inter;
  return reserve_payload(si);
}

bool HostCodec::prepare_frame(HostFrame& f, int* next_slot) {
  const bool more = prepare_frame_inner(f, next_slot);
  if (f.rc && f.msg.empty()) f.msg = msg;
  return more;
}

bool HostCodec::prepare_frame_inner(HostFrame& f, int* next_slot) {
  const uint8_t* data = f.host;
  const size_t len = f.len;
  if (len < 1) { f.rc = fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: empty frame payload"); return true; }
  const int type = data[0] & 3, coding = (data[0] >> 2) & 7;
  const bool gov = ((data[0] >> 5) & 1) != 0;
  if (type != 0 && type != 1) { f.rc = fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: frame type"); return true; }
  if (gov && type != 0) { f.rc = fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: golden override on a non-key frame"); return true; }
  size_t hlen = 1;
  if (type == 0) {
    if (len < 6) { f.rc = fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: key frame header truncated"); return true; }
    const int w = data[1] | data[2] << 8, h = data[3] | data[4] << 8;
    if (w == 0 || h == 0) { f.rc = fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: key frame dimensions"); return true; }
    if ((data[5] >> 4) != 0) { f.rc = fail(EBIV_ERR_UNSUPPORTED, "ebiv: unsupported stream feature: chroma format"); return true; }
    if ((data[5] & 15) != 0) { f.rc = fail(EBIV_ERR_UNSUPPORTED, "ebiv: unsupported stream feature: bit depth code"); return true; }
    hlen = 6;
    if ((f.rc = configure(w, h))) return true;
  } else if (!ready) {
    f.rc = fail(EBIV_ERR_NOKEYFRAME, "ebiv: inter frame before any key frame");
    return true;
  }
  f.ring_slot = next;
  if (++next == ring_size) next = 0;
  f.key = type == 0;
  if (coding == 0) {  // decodeRaw, codec.go:195-207
    const size_t want = static_cast<size_t>(W) * H + 2 * static_cast<size_t>(CW) * CH;
    if (len - hlen != want) { f.rc = fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: raw frame size"); return true; }
    f.raw = true;
    f.device_work = true;
    f.body_off = hlen;
    return true;
  }
  if (coding != 1 && coding != 2) { f.rc = fail(EBIV_ERR_UNSUPPORTED, "ebiv: unsupported stream feature: coding mode"); return true; }
  f.inter = coding == 2;
  const uint8_t* body = data + hlen;
  size_t blen = len - hlen;
  const uint8_t* ov = nullptr;
  size_t ovlen = 0;
  if (gov) {
    uint64_t kl = 0;
    const int k = uvarint(body, blen, &kl);
    if (k <= 0 || kl > static_cast<uint64_t>(blen - static_cast<size_t>(k))) {
      f.rc = fail(EBIV_ERR_CORRUPT, "ebiv: corrupt stream: golden override split");
      return true;
    }
    ov = body + k + kl;
    ovlen = blen - static_cast<size_t>(k) - static_cast<size_t>(kl);
    body += k;
    blen = static_cast<size_t>(kl);
  }
  if ((f.rc = prepare_coded(data, body, blen, f.inter, f.key, *next_slot))) return true;
  f.main_slot = (*next_slot)++;
  f.device_work = true;
  if (f.key && ov) {
    f.has_ov = true;
    f.saved_prev.assign(prev, prev + kNumContexts);
    f.ov_rc = prepare_coded(data, ov, ovlen, true, false, *next_slot);
    if (f.ov_rc) f.ov_msg = msg;
    else f.ov_slot = (*next_slot)++;
    return false;
  }
  return true;
}

int HostCodec::resolve(HostFrame& f, bool main_ok, bool ov_ok) {
  f.golden_written = false;
  if (f.rc) return f.rc;
  if (f.raw) return EBIV_OK;  // raw frames leave the references alone
  if (!main_ok) {
    // Go never reached the override, so never parsed its tables.
    if (f.has_ov) std::copy(f.saved_prev.begin(), f.saved_prev.end(), prev);
    return EBIV_ERR_CORRUPT;
  }
  if (f.has_ov) {
    if (f.ov_rc) {
      f.msg = f.ov_msg;
      return f.ov_rc;
    }
    f.golden_written = true;  // decoded, or partially rewritten on failure
    if (!ov_ok) {
      f.msg = f.ov_msg;
      return EBIV_ERR_CORRUPT;
    }
    return EBIV_OK;
  }
  if (f.key) f.golden_written = true;
  return EBIV_OK;
}

}  // namespace ebiv
