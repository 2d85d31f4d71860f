"""Writes the golden fixtures under tests/golden/: small EBIV streams made by
this repository's seeded generators and its C++ restatement of the govid
encoder, with the per-frame CRC-32 of each decoded plane (Y, Cb, Cr, packed
visible samples).

The expected CRCs come from the oracle (the C restatement of the Go
decoder) and are accepted only when they equal the encoder's own
reconstruction frame for frame, which is the format contract of the
reference (coded_test.go:107-124: decoder output == encoder
reconstruction).  No reference code runs here.

Run: python3 tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
import os
import sys
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

from govid_amd import clips, encoder  # noqa: E402

FIXTURES = {
    # the CPU-runnable plain drop-in case: 320x180, 30 frames, q22 gop30 1x1
    "cfg1_rects_320x180": dict(workload="rects", w=320, h=180, frames=30,
                               enc=dict(qp=22, gop=30, tile_cols=1, tile_rows=1, fast=False)),
    # a key frame carrying a golden override (alt-ref), two-pass
    "altref_parallax_64x48": dict(workload="parallax", w=64, h=48, frames=9,
                                  enc=dict(qp=28, gop=9, altref=4, fast=False)),
    # entropy-heavy: noise texture at q12, 2x2 tiles, short GOP
    "noise_q12_128x96_2x2": dict(workload="noisepan", w=128, h=96, frames=8,
                                 enc=dict(qp=12, gop=4, tile_cols=2, tile_rows=2)),
    # odd geometry: partial macroblocks, odd chroma, 2x3 tiles
    "odd_37x21_2x3": dict(workload="synthpan", w=37, h=21, frames=6,
                          enc=dict(qp=22, gop=6, tile_cols=2, tile_rows=3)),
    # raw (uncoded) frames
    "raw_48x32": dict(workload="synthframe", w=48, h=32, frames=3, enc=dict(coding="raw")),
}


def crcs(planes):
    return [zlib.crc32(bytes(memoryview(p.tobytes()))) for p in planes]


def make(name: str) -> dict:
    import oracle as O

    spec = FIXTURES[name]
    w, h, n = spec["w"], spec["h"], spec["frames"]
    seed = clips.stream_seed(spec["workload"], 0)
    frames = list(clips.generate(spec["workload"], w, h, n, seed))
    clip = encoder.encode_frames(frames, w, h, keep_recon=True, threads=4, **spec["enc"])
    dec = O.OracleCodec()
    out = []
    for i, pkt in enumerate(clip.packets):
        got = dec.decode(pkt)
        c = crcs(got)
        want = crcs(clip.recons[i]) if clip.recons else crcs(frames[i])
        if c != want:
            raise SystemExit(f"{name} frame {i}: oracle decode differs from the encoder reconstruction")
        out.append(c)
    with open(os.path.join(HERE, name + ".ebiv"), "wb") as fh:
        fh.write(clip.container)
    meta = {"name": name, "workload": spec["workload"], "width": w, "height": h, "frames": n, "seed": seed,
            "encoder": spec["enc"], "container_bytes": len(clip.container),
            "container_crc32": zlib.crc32(clip.container), "plane_crc32": out}
    with open(os.path.join(HERE, name + ".json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    return meta


if __name__ == "__main__":
    for nm in sys.argv[1:] or list(FIXTURES):
        m = make(nm)
        print(f"{nm}: {m['container_bytes']} bytes, {m['frames']} frames")
