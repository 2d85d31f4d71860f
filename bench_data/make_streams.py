"""Produces the benchmark's input streams with the package's own generator and
encoder, deterministically (seeded), so GPU runs need not re-encode them.

  parallax_1080p.ebiv — 1920x1080, 120 frames: three Perlin parallax layers
  (0.5 / 1 / 2 px per frame), 2-frame flash every 24 frames, noise sigma 1.5;
  q22, gop 60, 4x3 tiles, single-pass encode.
  parallax_360p / parallax_720p / parallax_2160p.ebiv — the serving mix's
  other sizes with the same content and encoding (640x360 and 1280x720 with
  4x3 tiles, 120 frames; 3840x2160 with 8x6 tiles, 60 frames).

Larger sets, written to bench_data/gen/ (kept out of version control; the
benchmark skips a section whose streams are absent):

  noisepan_1080p.ebiv — the entropy-bound workload: 1920x1080 mid-grey plus
  uniform noise (sigma 20) under random global motion in +-16 px, q12, gop 15,
  4x3 tiles; 30 frames (two GOPs), which the benchmark's 32 sessions replay
  from staggered key frames (about 1.2 MB per frame, so a full 120-frame
  stream per session would be 4.6 GB of input).
  serve_360p / serve_720p / serve_1080p / serve_2160p.ebiv — the serving
  mix: rects content, q22, gop 30, one tile (8x6 tiles at 3840x2160), 600
  frames each.
  mixed_00 .. mixed_63.ebiv — the mixed batch: 24 streams of 640x360, 24 of
  1280x720 and 16 of 1920x1080, rects content, each with its own draw of
  tile grid {1x1, 2x2, 4x3}, gop {30, 60, 120}, q {18, 22, 28} and length
  60-240 frames (seeded; see mixed_specs).

Run: python3 bench_data/make_streams.py [--jobs N] [--missing] [name ... | noisepan_1080p | mixed | serve]
"""
from __future__ import annotations

import json
import os
import sys
import time
import zlib

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


GEN = os.path.join(HERE, "gen")

STREAMS["noisepan_1080p"] = dict(workload="noisepan", w=1920, h=1080, frames=30, seed=0, dir=GEN,
                                 enc=dict(qp=12, gop=15, tile_cols=4, tile_rows=3))


def mixed_specs():
    """The 64 mixed-batch streams: sizes in the 24 / 24 / 16 split, each
    stream's tile grid, gop, q and length drawn from one seeded generator."""
    import numpy as np

    rng = np.random.default_rng(clips.stream_seed("mixed", 0))
    sizes = [(640, 360)] * 24 + [(1280, 720)] * 24 + [(1920, 1080)] * 16
    specs = {}
    for i, (w, h) in enumerate(sizes):
        tc, tr = ((1, 1), (2, 2), (4, 3))[int(rng.integers(3))]
        gop = (30, 60, 120)[int(rng.integers(3))]
        q = (18, 22, 28)[int(rng.integers(3))]
        frames = int(rng.integers(60, 241))
        specs[f"mixed_{i:02d}"] = dict(workload="rects", w=w, h=h, frames=frames, seed=100 + i, dir=GEN,
                                       enc=dict(qp=q, gop=gop, tile_cols=tc, tile_rows=tr))
    return specs


STREAMS.update(mixed_specs())


def serving_specs():
    """The serving mix's streams: rects content encoded like the 320x180
    clip (q22, gop 30, one tile) at 640x360, 1280x720 and 1920x1080, and
    3840x2160 with 8x6 tiles; 600 frames each, so that plays of 30-600
    frames need no clamping."""
    out = {}
    for w, h, tc, tr in ((640, 360, 1, 1), (1280, 720, 1, 1), (1920, 1080, 1, 1), (3840, 2160, 8, 6)):
        out[f"serve_{h}p"] = dict(workload="rects", w=w, h=h, frames=600, seed=200 + h, dir=GEN,
                                  enc=dict(qp=22, gop=30, tile_cols=tc, tile_rows=tr))
    return out


STREAMS.update(serving_specs())


def path_of(name: str) -> str:
    return os.path.join(STREAMS[name].get("dir", HERE), name + ".ebiv")


def crc_path_of(name: str) -> str:
    return path_of(name)[:-len(".ebiv")] + ".crc.json"


def make(name: str, threads: int = 8) -> str:
    """Encodes stream `name` and writes it with its expected plane CRCs: the
    CRC-32 of each frame's packed Y, Cb and Cr as the encoder reconstructed
    them (a decoder's output equals the encoder's reconstruction,
    ebiv/coded_test.go:107-124)."""
    spec = STREAMS[name]
    w, h, n = spec["w"], spec["h"], spec["frames"]
    seed = clips.stream_seed(spec["workload"], spec["seed"])
    enc = encoder.Encoder(w, h, threads=threads, keep_recon=True, **spec["enc"])
    t0 = time.time()
    for i, f in enumerate(clips.generate(spec["workload"], w, h, n, seed)):
        enc.write(*f)
        if i % 10 == 0:
            print(f"{name}: frame {i}/{n} {time.time() - t0:.0f}s", flush=True)
    clip = enc.close()
    path = path_of(name)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path + ".tmp", "wb") as fh:
        fh.write(clip.container)
    os.replace(path + ".tmp", path)
    crcs = [[zlib.crc32(p.tobytes()) for p in planes] for planes in clip.recons]
    with open(crc_path_of(name), "w") as fh:
        json.dump({"stream": name + ".ebiv", "stream_crc32": zlib.crc32(clip.container), "width": w, "height": h,
                   "source": "encoder reconstruction", "planes": "packed Y, Cb, Cr", "crc32": crcs}, fh)
        fh.write("\n")
    print(f"{name}: {len(clip.container)} bytes, {len(clip.packets)} frames -> {path}", flush=True)
    return path


def _make_one(name: str) -> str:
    # one thread per stream, four for the 3840x2160 one (48 tiles)
    return make(name, threads=4 if STREAMS[name]["w"] * STREAMS[name]["h"] > 4_000_000 else 1)


if __name__ == "__main__":
    args = sys.argv[1:]
    jobs = 1
    if args[:1] == ["--jobs"]:
        jobs, args = int(args[1]), args[2:]
    skip = args[:1] == ["--missing"]  # only streams whose file is absent
    if skip:
        args = args[1:]
    names = []
    for a in args or [k for k, v in STREAMS.items() if "dir" not in v]:
        if a in ("mixed", "serve"):
            names += sorted(k for k in STREAMS if k.startswith(a + "_"))
        else:
            names.append(a)
    if skip:
        names = [n for n in names if not os.path.exists(path_of(n))]
    if jobs > 1:
        from multiprocessing import Pool

        # longest encodes first so the pool drains evenly
        names.sort(key=lambda k: -STREAMS[k]["w"] * STREAMS[k]["h"] * STREAMS[k]["frames"])
        with Pool(jobs) as pool:
            for _ in pool.imap_unordered(_make_one, names):
                pass
    else:
        for nm in names:
            make(nm)
