"""Produces the benchmark's input streams with the package's own generator and
encoder, deterministically (seeded), so GPU runs need not re-encode them.

  parallax_1080p.ebiv — 1920x1080, 120 frames: three Perlin parallax layers
  (0.5 / 1 / 2 px per frame), 2-frame flash every 24 frames, noise sigma 1.5;
  q22, gop 60, 4x3 tiles, single-pass encode.
  parallax_360p / parallax_720p / parallax_2160p.ebiv — the serving mix's
  other sizes with the same content and encoding (640x360 and 1280x720 with
  4x3 tiles, 120 frames; 3840x2160 with 8x6 tiles, 60 frames).

Run: python3 bench_data/make_streams.py [name ...]
"""
from __future__ import annotations

import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from govid_amd import clips, encoder  # noqa: E402

STREAMS = {
    "parallax_1080p": dict(workload="parallax", w=1920, h=1080, frames=120, seed=0,
                           enc=dict(qp=22, gop=60, tile_cols=4, tile_rows=3)),
    "parallax_360p": dict(workload="parallax", w=640, h=360, frames=120, seed=0,
                          enc=dict(qp=22, gop=60, tile_cols=4, tile_rows=3)),
    "parallax_720p": dict(workload="parallax", w=1280, h=720, frames=120, seed=0,
                          enc=dict(qp=22, gop=60, tile_cols=4, tile_rows=3)),
    "parallax_2160p": dict(workload="parallax", w=3840, h=2160, frames=60, seed=0,
                           enc=dict(qp=22, gop=60, tile_cols=8, tile_rows=6)),
}


def make(name: str, threads: int = 8) -> str:
    spec = STREAMS[name]
    w, h, n = spec["w"], spec["h"], spec["frames"]
    seed = clips.stream_seed(spec["workload"], spec["seed"])
    enc = encoder.Encoder(w, h, threads=threads, **spec["enc"])
    t0 = time.time()
    for i, f in enumerate(clips.generate(spec["workload"], w, h, n, seed)):
        enc.write(*f)
        if i % 10 == 0:
            print(f"{name}: frame {i}/{n} {time.time() - t0:.0f}s", flush=True)
    clip = enc.close()
    path = os.path.join(HERE, name + ".ebiv")
    with open(path, "wb") as fh:
        fh.write(clip.container)
    print(f"{name}: {len(clip.container)} bytes, {len(clip.packets)} frames -> {path}", flush=True)
    return path


if __name__ == "__main__":
    for nm in sys.argv[1:] or list(STREAMS):
        make(nm)
