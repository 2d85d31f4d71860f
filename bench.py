"""EBIV decode benchmark on one MI355X.

Workload: one 1920x1080 4:2:0 clip, 120 frames (three Perlin parallax layers,
2-frame flash every 24 frames, noise sigma 1.5), EBIV q22, gop 60, 4x3 tiles.

  plain    the drop-in: one Codec, one packet per Codec.Decode call (batch 1).
           A step is one pass over the clip.  This is `value`.
  batched  the same clip as S concurrent sessions of one Batch engine, each
           starting at a different key frame and wrapping round (a loop
           restart), decoding up to W frames per session per engine step.
           A step is one pass of every session over the clip.
  mixed    64 sessions, one stream each: 24 x 640x360, 24 x 1280x720 and
           16 x 1920x1080 (rects content), each with its own tile grid
           {1x1, 2x2, 4x3}, gop {30, 60, 120}, q {18, 22, 28} and length
           60-240 frames.  A step is one pass of every session over its stream.
  entropy  32 sessions of 1920x1080 noise (sigma 20, random global motion in
           +-16 px), q12, gop 15, 4x3 tiles: the symbol-heavy case, reported
           in Msymbols/s.  The sessions replay one 30-frame stream from
           staggered key frames.
  serving  Poisson arrivals of play / seek / loop sessions of 360p-2160p
           streams under an HBM budget; p50/p99 frame latency.

The mixed and entropy streams are generated by bench_data/make_streams.py
into bench_data/gen/; a section whose streams are absent is reported as
skipped.  Packets are resident in HBM before the timed regions and frames
stay in HBM.  Every frame of the plain pass and of each batched warmup pass,
and a sample of the served frames, is checked bit for bit (plane CRCs)
against the expected CRCs stored beside each stream, outside the timed
regions; those are the encoder's reconstruction, which the CPU tests check
against the oracle decoder.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
import zlib

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` in `workload` from the committed
    rocprofv3 counter summary (profiles/hbm_traffic.json): FETCH_SIZE and
    WRITE_SIZE in KiB per dispatch, each scaled by what the same counter
    reads for a plain device copy of known size.  None when not measured."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as fh:
            t = json.load(fh)
        k = t["workloads"][workload][kernel]
        cal = t["calibration"]
        return round(1024 * (k["fetch_kib"] * cal["fetch_scale"] + k["write_kib"] * cal["write_scale"]))
    except (OSError, KeyError, ValueError):
        return None
BENCH_DATA = os.path.join(ROOT, "bench_data")
GEN = os.path.join(BENCH_DATA, "gen")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stream_path(name, gen=False):
    return os.path.join(GEN if gen else BENCH_DATA, name + ".ebiv")


def load_stream(name="parallax_1080p"):
    """Reader and expected plane CRCs of a committed benchmark stream."""
    from govid_amd.container import Reader

    path = stream_path(name)
    if not os.path.exists(path):
        log(f"bench: {name} missing, generating it (several minutes)")
        sys.path.insert(0, BENCH_DATA)
        import make_streams

        make_streams.make(name, threads=os.cpu_count() or 8)
    rd = Reader.open(path)
    return rd, expected_crcs(path, rd.data)


def expected_crcs(path, data):
    """Per-frame (Y, Cb, Cr) CRC-32s stored beside a stream by
    bench_data/make_streams.py (the encoder's reconstruction)."""
    with open(path[:-len(".ebiv")] + ".crc.json") as fh:
        d = json.load(fh)
    if d["stream_crc32"] != zlib.crc32(data):
        raise RuntimeError(f"{path}: the CRC file belongs to another version of the stream")
    return [tuple(c) for c in d["crc32"]]


def _crc(planes):
    return tuple(zlib.crc32(p.tobytes()) for p in planes)


def cpu_baseline(packets, budget_s=20.0):
    """The oracle on this host: tile-parallel like ebiv/pool.go."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = min(os.cpu_count() or 1, 16)
    dec = O.OracleCodec(threads=cores)
    t0 = time.perf_counter()
    n = 0
    while n < len(packets):
        dec.decode(packets[n])
        n += 1
        if time.perf_counter() - t0 > budget_s and n >= 8:
            break
    dt = time.perf_counter() - t0
    return {"value": round(n / dt, 2), "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"first {n} of {len(packets)} frames of the same 1080p clip (C restatement of the "
                      f"govid ebiv tiled decoder, one tile per job over {cores} threads)"}


def verify_plain(packets, want):
    from govid_amd.codec import Codec

    gpu = Codec(ring=3)
    bad = sum(1 for p, w in zip(packets, want) if _crc(gpu.decode(p).planes()) != w)
    gpu.close()
    return bad


def pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(q * len(v)))]


def run_plain(packets, dev, steps, warmup):
    from govid_amd.codec import Codec

    codec = Codec(ring=3)
    for _ in range(warmup):
        for p, d in zip(packets, dev):
            codec.decode_device(p, d.ptr)
    s0 = codec.stats()
    codec.set_timing(True)
    lat = []
    t0 = time.perf_counter()
    for step in range(steps):
        for p, d in zip(packets, dev):
            a = time.perf_counter()
            codec.decode_device(p, d.ptr)
            lat.append(time.perf_counter() - a)
        log(f"bench: plain step {step} done")
    elapsed = time.perf_counter() - t0
    kt = codec.kernel_times()
    s1 = codec.stats()
    codec.close()
    return {"elapsed": elapsed, "frames": steps * len(packets), "lat": lat, "kt": kt,
            "symbols": s1["symbols"] - s0["symbols"]}


def run_batched(packets, dev, want, sessions, window, steps, warmup):
    """S sessions over one clip, each from a different key frame."""
    keys = [i for i, p in enumerate(packets) if p[0] & 3 == 0]
    plan = []
    for k in range(sessions):
        st = keys[k % len(keys)]
        plan.append((packets, dev, want, list(range(st, len(packets))) + list(range(0, st))))
    return run_sessions(plan, window, steps, warmup)


def run_sessions(plan, window, steps, warmup):
    """One Batch engine; plan = [(packets, device buffers, expected CRCs,
    frame order)] per session.  A pass submits every session's frames in
    its order and steps the engine until all have completed; the first
    warmup pass checks every frame."""
    from govid_amd.codec import Batch

    b = Batch()
    sids = [b.open(ring=window, window=window) for _ in plan]

    def one_pass(check):
        for sid, (packets, dev, _, order) in zip(sids, plan):
            b.submit_many(sid, [packets[i] for i in order], [dev[i].ptr for i in order], tags=order)
        wants = {sid: w for sid, (_, _, w, _) in zip(sids, plan)}
        lat, bad, n = [], 0, 0
        while b.pending():
            a = time.perf_counter()
            done = b.step()
            lat.append(time.perf_counter() - a)
            for c in done:
                n += 1
                if c.status:
                    bad += 1
                elif check and _crc(c.frame.planes()) != wants[c.session][c.tag]:
                    bad += 1
        return lat, bad, n

    bad = 0
    for w in range(max(warmup, 1)):
        _, bw, _ = one_pass(check=(w == 0))
        bad += bw
    s0 = b.stats()
    b.set_timing(True)
    lat, frames = [], 0
    t0 = time.perf_counter()
    for step in range(steps):
        l, bw, n = one_pass(check=False)
        lat += l
        bad += bw
        frames += n
        log(f"bench: batched step {step} done ({n} frames)")
    elapsed = time.perf_counter() - t0
    kt = b.kernel_times()
    s1 = b.stats()
    b.close()
    return {"elapsed": elapsed, "frames": frames, "lat": lat, "kt": kt, "bad": bad,
            "symbols": s1["symbols"] - s0["symbols"]}


def batched_summary(bt, steps, **extra):
    bk = bt["kt"]
    return dict(extra, **{
        "frames_per_s": round(bt["frames"] / bt["elapsed"], 2),
        "frames_per_step": bt["frames"] // steps,
        "engine_step_p50_ms": round(statistics.median(bt["lat"]) * 1e3, 3),
        "engine_step_p99_ms": round(pct(bt["lat"], 0.99) * 1e3, 3),
        "msymbols_per_s": round(bt["symbols"] / bt["elapsed"] / 1e6, 2),
        "frames_differing": bt["bad"],
        "kernel_avg_us": {k: round(kavg_us(bk, k), 2) for k in ("parse", "inter", "intra")},
        "kernel_launches": {k: bk[f"{k}_n"] for k in ("parse", "inter", "intra")},
        "recon": {"ebiv_recon_inter": inter_roofline(bk) or {}},
    })


def open_gen_streams(names):
    """Readers, expected CRCs and device buffers of generated streams, or
    None when any is absent."""
    from govid_amd.codec import DeviceBuffer
    from govid_amd.container import Reader

    out = []
    for nm in names:
        path = stream_path(nm, gen=True)
        if not os.path.exists(path):
            return None
        rd = Reader.open(path)
        pk = rd.packets()
        out.append((nm, rd, pk, [DeviceBuffer(p) for p in pk], expected_crcs(path, rd.data)))
    return out


def run_mixed(window, steps, warmup):
    """The mixed batch: 64 sessions, one generated stream each."""
    names = [f"mixed_{i:02d}" for i in range(64)]
    st = open_gen_streams(names)
    if st is None:
        return {"skipped": "streams absent: python3 bench_data/make_streams.py --jobs 8 mixed"}
    plan = [(pk, dev, want, list(range(len(pk)))) for _, _, pk, dev, want in st]
    bt = run_sessions(plan, window, steps, warmup)
    for *_, dev, _ in st:
        for d in dev:
            d.close()
    sizes = {}
    for _, rd, pk, _, _ in st:
        k = f"{rd.header.width}x{rd.header.height}"
        sizes[k] = sizes.get(k, 0) + 1
    hd = sum(len(pk) for _, rd, pk, _, _ in st if rd.header.width == 1920) * steps
    out = batched_summary(bt, steps, sessions=len(st), window=window, streams=sizes,
                          frames_per_pass=sum(len(pk) for _, _, pk, _, _ in st))
    out["frames_1080p_share"] = round(hd / max(bt["frames"], 1), 3)
    return out


def run_entropy(sessions, window, steps, warmup):
    """The entropy-bound batch: S sessions of the 1080p q12 noise stream."""
    st = open_gen_streams(["noisepan_1080p"])
    if st is None:
        return {"skipped": "stream absent: python3 bench_data/make_streams.py noisepan_1080p"}
    _, rd, pk, dev, want = st[0]
    bt = run_batched(pk, dev, want, sessions, window, steps, warmup)
    for d in dev:
        d.close()
    out = batched_summary(bt, steps, sessions=sessions, window=window,
                          stream=f"1920x1080 noise q12 gop15 4x3 tiles, {len(pk)} frames, "
                                 f"{sum(map(len, pk)) // len(pk)} B per frame")
    out["symbols_per_frame"] = round(bt["symbols"] / max(bt["frames"], 1))
    return out


SERVE_STREAMS = ("parallax_360p", "parallax_720p", "parallax_1080p", "parallax_2160p")


def run_serving(requests_n, budget_bytes, window, check_every, seed=5):
    """The serving mix: Poisson arrivals (200 requests/s) over 360p, 720p and
    1080p streams (4x3 tiles) and a 2160p stream (8x6 tiles), all of the
    parallax content at q22, gop 60; 20% seek to a random key frame part-way, 10% loop
    restarts; lengths 30-600 frames; sessions admitted under an HBM budget.
    Every `check_every`-th completed frame is compared with the stream's
    expected CRCs for that frame."""
    from govid_amd.codec import Batch, DeviceBuffer
    from govid_amd.serve import Server, Stream, poisson_requests

    streams, bufs, want = [], [], {}
    for nm in SERVE_STREAMS:
        rd, want[nm] = load_stream(nm)
        pk = rd.packets()
        dev = [DeviceBuffer(p) for p in pk]
        bufs += dev
        streams.append(Stream(nm, rd.header.width, rd.header.height, pk, [d.ptr for d in dev]))
    reqs = poisson_requests(streams, requests_n, rate_per_s=200.0, seed=seed, min_len=30, max_len=600)
    checked = [0, 0]
    count = [0]

    def on_frame(fd):
        count[0] += 1
        if fd.status == 0 and count[0] % check_every == 0:
            checked[0] += 1
            st = next(r.stream for r in reqs if r.rid == fd.rid)
            if _crc(fd.frame.planes()) != want[st.name][fd.index]:
                checked[1] += 1

    b = Batch()
    srv = Server(b, budget_bytes, window=window, ring=window, on_frame=on_frame)
    stats = srv.run(reqs, max_seconds=240)
    b.close()
    for d in bufs:
        d.close()
    stats.update({"streams": [f"{s.name} {s.width}x{s.height} x{len(s.packets)}" for s in streams],
                  "budget_bytes": budget_bytes, "window": window,
                  "frames_checked": checked[0], "frames_differing": checked[1]})
    return stats


def kavg_us(kt, k):
    return kt[f"{k}_ms"] / max(kt[f"{k}_n"], 1) * 1e3


# Inter reconstruction moves, per macroblock, its 384 predicted samples
# (256 luma + 2 x 64 chroma) in from the reference and 384 reconstructed
# samples out; filter margins re-read from cache are not algorithmic bytes.
INTER_BYTES_PER_MB = 768


def inter_roofline(kt):
    n, t = kt["inter_n"], kt["inter_ms"] / 1e3
    if not n or not t:
        return None
    per_launch = INTER_BYTES_PER_MB * kt["inter_mbs"] / n
    return {"bytes_per_launch": round(per_launch), "macroblocks_per_launch": round(kt["inter_mbs"] / n, 1),
            "avg_launch_us": round(t / n * 1e6, 2), "achieved_GBps": round(per_launch * n / t / 1e9, 2),
            "frac": round(per_launch * n / t / HBM_PEAK, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sessions", type=int, default=64)
    # frames per session per engine step in the batched sections: a wider
    # window puts more independent tile chains into each entropy launch and
    # spreads the per-step costs over more frames (8 / 16 / 32 on the 64 x 1080p
    # clip: 37.6k / 38.4k / 40.2k frames/s; on the mixed batch 27.5k / 35.4k /
    # 41.0k), at one step's latency per frame
    ap.add_argument("--window", type=int, default=32)
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-only", action="store_true", help="skip the plain pass (for profiling the batched one)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-entropy", action="store_true")
    ap.add_argument("--entropy-sessions", type=int, default=32)
    ap.add_argument("--serve-only", action="store_true", help="run only the serving mix")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-budget-gb", type=float, default=24.0)
    ap.add_argument("--serve-window", type=int, default=1,
                    help="frames per session per engine step when serving (1 keeps p99 frame latency lowest)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 2 GiB device-to-device copy (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        # 2 GiB per buffer, well above the 256 MB last-level cache, so each
        # copy streams from and to HBM
        n = 2 << 30
        x = torch.randint(0, 255, (n,), dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy_calibration", "bytes_read": n, "bytes_written": n, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this decoder runs on one GPU; runs on several GPUs are not part of this "
                                   "build (sessions would shard by stream across GPUs)"}))
        sys.exit(2)

    from govid_amd.codec import DeviceBuffer

    if args.serve_only:
        sv = run_serving(args.serve_requests, int(args.serve_budget_gb * (1 << 30)), args.serve_window, check_every=17)
        print(json.dumps({"mode": "serving", **{k: (round(v, 3) if isinstance(v, float) else v) for k, v in sv.items()}}))
        return
    rd, want = load_stream()
    packets = rd.packets()
    hdr = rd.header
    W, H = hdr.width, hdr.height
    log(f"bench: {len(packets)} packets, {sum(map(len, packets))} bytes, {W}x{H}")

    dev = [DeviceBuffer(p) for p in packets]
    if args.batched_only:
        bt = run_batched(packets, dev, want, args.sessions, args.window, args.steps, args.warmup)
        print(json.dumps({"mode": "batched", "frames_per_s": round(bt["frames"] / bt["elapsed"], 2),
                          "frames_differing": bt["bad"], "sessions": args.sessions, "window": args.window}))
        return
    bad = 0 if args.no_verify else verify_plain(packets, want)
    log(f"bench: plain verify -> {bad} frames differ from the expected CRCs")

    pl = run_plain(packets, dev, args.steps, args.warmup)
    fps = pl["frames"] / pl["elapsed"]
    kt = pl["kt"]
    log(f"bench: plain {fps:.1f} frames/s")

    # Dominant kernel of the plain path: the entropy kernel.  Algorithmic
    # bytes per launch: the frame's coded bytes read plus its macroblock
    # records written (48 B per macroblock), averaged over the clip; the
    # coefficient entries it also writes (4 B per nonzero level) are left out.
    nmb = ((W + 15) // 16) * ((H + 15) // 16)
    coded = sum(map(len, packets)) / len(packets)
    parse_bytes = coded + 48 * nmb
    parse_avg_s = kavg_us(kt, "parse") / 1e6
    roof = {"bound": "hbm", "kernel": "ebiv_parse", "achieved": round(parse_bytes / parse_avg_s / 1e9, 3),
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(parse_bytes / parse_avg_s / HBM_PEAK, 6),
            "traffic": measured_traffic("plain_1080p_parallax", "ebiv_parse"), "bytes_per_launch": round(parse_bytes), "avg_launch_us": round(parse_avg_s * 1e6, 2),
            "note": "latency-bound serial rANS chains (one wavefront per tile); reported in Msymbols/s"}
    recon = {"ebiv_recon_inter": dict(inter_roofline(kt) or {},
                                      traffic=measured_traffic("plain_1080p_parallax", "ebiv_recon_inter")),
             "ebiv_recon_intra": {"avg_launch_us": round(kavg_us(kt, "intra"), 2)}}

    out = {
        "metric": "EBIV 1080p decoded frames/s bit-exact, batched vs plain GPU; p50/p99 frame ms",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(pl["elapsed"] / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "plain drop-in, 1920x1080 parallax clip, 120 frames, q22 gop60 4x3 tiles, batch 1",
                   "frames_per_step": len(packets), "mode": "plain"},
        "p50_frame_ms": round(statistics.median(pl["lat"]) * 1e3, 3),
        "p99_frame_ms": round(pct(pl["lat"], 0.99) * 1e3, 3),
        "msymbols_per_s": round(pl["symbols"] / pl["elapsed"] / 1e6, 2),
        "bitexact_frames_differing": bad,
        "kernel_avg_us": {k: round(kavg_us(kt, k), 2) for k in ("parse", "inter", "intra")},
        "roofline": roof,
        "recon": recon,
    }
    if not args.no_batched:
        bt = run_batched(packets, dev, want, args.sessions, args.window, args.steps, args.warmup)
        bfps = bt["frames"] / bt["elapsed"]
        log(f"bench: batched {bfps:.1f} frames/s")
        out["batched"] = batched_summary(bt, args.steps, vs_plain=round(bfps / fps, 2),
                                         sessions=args.sessions, window=args.window)
        out["batched"]["recon"]["ebiv_recon_inter"]["traffic"] = measured_traffic(
            "batched_64x1080p_parallax", "ebiv_recon_inter")
    for d in dev:
        d.close()
    if not args.no_mixed:
        out["mixed"] = run_mixed(args.window, args.steps, args.warmup)
        log(f"bench: mixed {out['mixed'].get('frames_per_s', out['mixed'].get('skipped'))}")
    if not args.no_entropy:
        out["entropy"] = run_entropy(args.entropy_sessions, args.window, args.steps, args.warmup)
        log(f"bench: entropy {out['entropy'].get('msymbols_per_s', out['entropy'].get('skipped'))} Msymbols/s")
    if not args.no_serve:
        sv = run_serving(args.serve_requests, int(args.serve_budget_gb * (1 << 30)), args.serve_window, check_every=17)
        log(f"bench: serving {sv['frames_per_s']:.1f} frames/s, p99 {sv['p99_frame_ms']:.2f} ms")
        out["serving"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in sv.items()}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(packets)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
