// host_replay.cpp — test driver (not part of the library): replays a packet
// sequence through the decoder's host half (govid_amd/csrc/ebiv_host.cpp, the
// same code libebivgpu.so runs) and prints each packet's status.  Where the
// device would decode tiles, the tile outcome is read from the input (the
// test takes it from the oracle).  Built by tests/test_host_replay.py with
// AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Input file: u32 ring size, then records of
//   u8 op (0 decode, 1 flush, 2 a new codec), and for decode: u32 length, the packet bytes,
//   u8 main payload tiles decoded (1/0), u8 override tiles decoded (1/0).
// Output: one line per decode, "<status> <ring slot or -1>".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ebiv_host.h"

namespace {

bool read_exact(FILE* f, void* dst, size_t n) { return n == 0 || fread(dst, 1, n, f) == n; }

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: host_replay FILE\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  uint32_t ring = 0;
  if (!read_exact(f, &ring, 4)) return 2;
  const int ring_size = ring >= 2 ? static_cast<int>(ring) : 3;
  ebiv::HostCodec* codec = new ebiv::HostCodec();
  codec->ring_size = ring_size;
  for (;;) {
    uint8_t op = 0;
    if (!read_exact(f, &op, 1)) break;
    if (op == 1) {  // Codec.Flush (codec.go:115-117)
      codec->next = 0;
      continue;
    }
    if (op == 2) {  // NewCodec
      delete codec;
      codec = new ebiv::HostCodec();
      codec->ring_size = ring_size;
      continue;
    }
    uint32_t len = 0;
    if (!read_exact(f, &len, 4)) return 2;
    // An exact-size heap copy, so the sanitizer sees any read past the packet.
    uint8_t* pkt = static_cast<uint8_t*>(malloc(len ? len : 1));
    uint8_t ok[2] = {1, 1};
    if (!read_exact(f, pkt, len) || !read_exact(f, ok, 2)) return 2;
    ebiv::HostFrame fr;
    fr.host = pkt;
    fr.len = len;
    int slot = 0;
    codec->prepare_frame(fr, &slot);
    const int st = codec->resolve(fr, ok[0] != 0, ok[1] != 0);
    printf("%d %d\n", st, st == 0 ? fr.ring_slot : -1);
    free(pkt);
  }
  delete codec;
  fclose(f);
  return 0;
}
