"""Launch-shape tuner for the batch engine.

The batch engine's intra reconstruction runs one workgroup per tile with a
configurable number of wavefronts (ebiv_batch_set_schedule): key-frame tiles
hold every macroblock of the tile as intra work, inter-frame tiles only a
few, and a launch of many tiles fills the chip with small workgroups better
than with wide ones.  The best shapes depend on the caller's mix (tile size,
key-frame interval, sessions per step), so this module measures candidate
schedules on the caller's own stream and reports the fastest by wall time
per frame (the engine overlaps launches on several streams, so summed kernel
time is not the measure).  Every
candidate's decoded frames are compared with the first candidate's (frame
CRCs per packet index): schedules must not change results.

The window (frames per session per engine step) is searched first when
several are given: a wider one puts more independent tile chains into each
entropy launch and spreads each step's fixed costs over more frames, at one
step's latency per frame and (ring + window) frame buffers per session.  It
is picked by wall time per frame, and so is the schedule.

    python -m govid_amd.tune STREAM.ebiv [--sessions 64] [--window 4 | --windows 8,16,32] [--passes 2]

prints one JSON object: the windows and schedules tried with their time per
frame, and the fastest.
"""
from __future__ import annotations

import argparse
import itertools
import json
import time
import zlib

from .codec import Batch, DeviceBuffer
from .container import Reader

# Candidate values; the engine accepts 1-16 wavefronts per tile.
KEY_WAVES = (8, 16)
KEY_MANY_WAVES = (2, 4, 8)
FEW_WAVES = (4, 8, 16)
MANY_WAVES = (2, 4, 8)
MANY_TILES = (128, 256, 512)
RECON_LANES = (1, 2, 4)
PIPELINE = (1, 2)  # next step's entropy launch beside this step's reconstruction, or not
KEY_BAND_ROWS = (2, 4, 8, 128)  # rows per workgroup of a key-frame tile (128: whole tiles)


def _crc(planes) -> int:
    c = 0
    for p in planes:
        c = zlib.crc32(p.tobytes(), c)
    return c


def measure(packets, dev, schedule: dict, sessions: int, window: int, passes: int, check: bool):
    """Kernel milliseconds per frame of `schedule` over `passes` passes of
    every session through the whole stream (sessions start at staggered key
    frames), and the frame CRCs of the first pass when `check` is set."""
    keys = [i for i, p in enumerate(packets) if p[0] & 3 == 0] or [0]
    b = Batch()
    try:
        b.set_schedule(**schedule)
        sids = [b.open(ring=window, window=window) for _ in range(sessions)]
        orders = []
        for k in range(sessions):
            st = keys[k % len(keys)]
            orders.append(list(range(st, len(packets))) + list(range(0, st)))
        crcs: dict = {}

        def one_pass(record):
            for sid, order in zip(sids, orders):
                b.submit_many(sid, [packets[i] for i in order], [dev[i].ptr for i in order], tags=order)
            n = 0
            while b.pending():
                for c in b.step():
                    n += 1
                    if c.status:
                        raise c.error
                    if record:
                        crcs.setdefault(c.tag, _crc(c.frame.planes()))
            return n

        one_pass(check)  # warm-up, and the results to compare
        b.set_timing(True)
        frames = 0
        t0 = time.perf_counter()
        for _ in range(passes):
            frames += one_pass(False)
        wall = time.perf_counter() - t0
        kt = b.kernel_times()
    finally:
        b.close()
    kernel_ms = sum(v for k, v in kt.items() if k.endswith("_ms"))
    return {"kernel_ms_per_frame": kernel_ms / max(frames, 1), "wall_ms_per_frame": 1e3 * wall / max(frames, 1),
            "frames": frames}, crcs


def tune(source, sessions: int = 64, window: int = 4, passes: int = 2, exhaustive: bool = False,
         windows=None) -> dict:
    """`source`: a container path or a list of packets; `windows`: candidate
    windows (searched first, under the engine's default schedule)."""
    packets = Reader.open(source).packets() if isinstance(source, str) else list(source)
    dev = [DeviceBuffer(p) for p in packets]
    window_results, ref = [], None
    if windows:
        base = Batch()
        d0 = base.schedule()
        base.close()
        for w in windows:
            r, crcs = measure(packets, dev, d0, sessions, w, passes, check=True)
            if ref is None:
                ref = crcs
            elif crcs != ref:
                raise RuntimeError(f"window {w} changed decoded frames")
            window_results.append(dict(window=w, **r))
        window = min(window_results, key=lambda r: r["wall_ms_per_frame"])["window"]
    if exhaustive:
        cands = [dict(intra_waves_key=a, intra_waves_key_many=k, intra_waves_few=b, intra_waves_many=c, many_tiles=d,
                      recon_lanes=e, pipeline=p)
                 for a, k, b, c, d, e, p in itertools.product(KEY_WAVES, KEY_MANY_WAVES, FEW_WAVES, MANY_WAVES,
                                                              MANY_TILES, RECON_LANES, PIPELINE)]
    else:  # one field at a time around the engine's defaults
        base = Batch()
        d0 = base.schedule()
        base.close()
        cands = [dict(d0)]
        for field, vals in (("intra_waves_key", KEY_WAVES), ("intra_waves_key_many", KEY_MANY_WAVES),
                            ("intra_waves_few", FEW_WAVES), ("intra_waves_many", MANY_WAVES),
                            ("many_tiles", MANY_TILES), ("recon_lanes", RECON_LANES), ("pipeline", PIPELINE),
                            ("key_band_rows", KEY_BAND_ROWS)):
            for v in vals:
                c = dict(d0, **{field: v})
                if c not in cands:
                    cands.append(c)
    results = []
    for c in cands:
        r, crcs = measure(packets, dev, c, sessions, window, passes, check=True)
        if ref is None:
            ref = crcs
        elif crcs != ref:
            raise RuntimeError(f"schedule {c} changed decoded frames")
        results.append(dict(schedule=c, **r))
    for d in dev:
        d.close()
    best = min(results, key=lambda r: r["wall_ms_per_frame"])
    name = source if isinstance(source, str) else f"{len(packets)} packets"
    return {"stream": name, "sessions": sessions, "window": window, "windows_tried": window_results,
            "tried": results, "best": best["schedule"]}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("stream")
    ap.add_argument("--sessions", type=int, default=64)
    ap.add_argument("--window", type=int, default=4)
    ap.add_argument("--windows", default="", help="comma-separated candidate windows (overrides --window)")
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--exhaustive", action="store_true", help="every combination instead of one field at a time")
    a = ap.parse_args(argv)
    wins = [int(w) for w in a.windows.split(",") if w] or None
    print(json.dumps(tune(a.stream, a.sessions, a.window, a.passes, a.exhaustive, windows=wins)))


if __name__ == "__main__":
    main()
