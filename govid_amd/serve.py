"""Continuous-batching decode server over the batched engine.

govid plays one stream per Codec from a frame source (source.go:9-33): `next`
decodes the next packet, `seek` repositions the demuxer and flushes the codec
(source.go:45-50, 226-234), and a looping player restarts at the first frame
(player.go:161-170).  Here many such playbacks share one GPU:

- A *request* names a stream and what to play: from the start, from a key frame
  (a seek), or the whole stream twice (a loop restart).
- Requests are admitted as sessions while the device memory budget allows: a
  session holds `session_bytes(w, h, ring, window)` of HBM for its frame pool
  and entropy buffers, and it keeps its references, golden frame and table
  history on the device for its whole life.  Requests that do not fit wait in
  arrival order and are admitted as running sessions finish and free memory.
- Every engine step decodes up to `window` queued frames of every active
  session together (one entropy launch, one reconstruction launch per frame
  position), so newly admitted sessions join the very next step: continuous
  batching, not a queue in front of a single decoder.

Per frame the server records the latency from the moment its packet was
submitted to the moment its step returned; per request, the admission wait
from its arrival to its admission, and the time from its arrival to its first
frame.  It checks nothing itself; callers compare frames (the tests do,
against the oracle).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field


@dataclass
class Stream:
    """A demuxed stream: its packets in order and its key-frame indices."""
    name: str
    width: int
    height: int
    packets: list
    dev_ptrs: list | None = None  # the same packets resident in device memory
    keys: list = field(default_factory=list)

    def __post_init__(self):
        if not self.keys:
            self.keys = [i for i, p in enumerate(self.packets) if p and (p[0] & 3) == 0]


@dataclass
class Request:
    """Play `stream` from frame `start` (a key frame) for `count` frames; past
    the last frame the playback restarts at frame 0 (a loop).  With
    `seek_at`, after that many frames the playback seeks to key frame
    `seek_key` and continues from there."""
    rid: int
    stream: Stream
    start: int = 0
    count: int | None = None
    arrival: float = 0.0  # seconds after the server starts
    seek_at: int | None = None
    seek_key: int = 0

    def frame_order(self):
        n = len(self.stream.packets)
        count = self.count if self.count is not None else n - self.start
        out, i = [], self.start
        for k in range(max(count, 1)):
            if self.seek_at is not None and k == self.seek_at:
                i = self.seek_key
            out.append(i)
            i += 1
            if i == n:
                i = 0
        return out


@dataclass
class FrameDone:
    rid: int
    index: int        # frame index in the stream
    seq: int          # position in the request's playback
    status: int
    latency_s: float
    frame: object     # engine frame (valid until the session decodes `ring` more)


class Server:
    """Admits requests under an HBM budget and runs the batched engine.

    `engine` is a govid_amd.codec.Batch (or any object with its interface:
    open, submit, flush, step, pending, close_session)."""

    def __init__(self, engine, hbm_budget: int, window: int = 4, ring: int = 4, session_bytes=None,
                 on_frame=None, clock=time.perf_counter):
        self.engine = engine
        self.budget = hbm_budget
        self.window = window
        self.ring = max(ring, window)
        self.on_frame = on_frame
        self.clock = clock
        if session_bytes is None:
            from .codec import session_bytes as sb
            session_bytes = sb
        self.session_bytes = session_bytes
        self.used = 0
        self.waiting: deque = deque()
        self.active: dict = {}   # session id -> _Run
        self.done: list = []
        self.frames: list = []
        self.max_active = 0
        self.peak_bytes = 0
        self.waits: list = []        # admission wait per request, seconds
        self.first_frame: list = []  # arrival to first completed frame, seconds
        self.step_times: list = []   # wall time of each engine step, seconds
        self.feed_times: list = []   # admission and packet submission before each step, seconds

    class _Run:
        def __init__(self, req, sid, nbytes, order):
            self.req, self.sid, self.nbytes = req, sid, nbytes
            self.order = order
            self.first_done = False
            self.next = 0          # next playback position to submit
            self.completed = 0
            self.submit_time = {}  # playback position -> submit time

    def _cost(self, req):
        return self.session_bytes(req.stream.width, req.stream.height, self.ring, self.window)

    def _admit(self, now):
        while self.waiting:
            req = self.waiting[0]
            cost = self._cost(req)
            if cost > self.budget:
                raise ValueError(f"request {req.rid} needs {cost} bytes, more than the whole budget")
            if self.used + cost > self.budget:
                break
            self.waiting.popleft()
            self.waits.append(max(0.0, now - req.arrival))
            sid = self.engine.open(ring=self.ring, window=self.window)
            self.used += cost
            self.peak_bytes = max(self.peak_bytes, self.used)
            self.active[sid] = self._Run(req, sid, cost, req.frame_order())
        self.max_active = max(self.max_active, len(self.active))

    def _feed(self, now):
        for run in self.active.values():
            # keep `window` packets of each session queued ahead
            while run.next < len(run.order) and run.next - run.completed < self.window:
                idx = run.order[run.next]
                st = run.req.stream
                if run.next > 0 and idx != run.order[run.next - 1] + 1:
                    # a jump (loop restart or seek): Codec.Flush, as
                    # source.go:45-50 does after Seek.  Flush drops queued
                    # packets, so it waits until the session has drained.
                    if run.completed < run.next:
                        break
                    self.engine.flush(run.sid)
                dev = st.dev_ptrs[idx] if st.dev_ptrs else None
                self.engine.submit(run.sid, st.packets[idx], dev, tag=run.next)
                run.submit_time[run.next] = now
                run.next += 1

    def run(self, requests, max_seconds: float = 600.0, progress=None, progress_every: float = 5.0):
        """Serves `requests` (sorted by arrival) to completion; returns stats.
        `progress`, when given, is called with a one-line status every
        `progress_every` seconds."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = self.clock()
        k = 0
        steps = 0
        next_report = progress_every
        while k < len(pending) or self.waiting or self.active:
            now = self.clock() - t0
            if progress and now >= next_report:
                next_report += progress_every
                progress(f"{now:.0f}s: {len(self.frames)} frames, {len(self.active)} active, "
                         f"{len(self.waiting)} waiting, {len(pending) - k} to arrive")
            if now > max_seconds:
                raise TimeoutError("serving did not finish in time")
            while k < len(pending) and pending[k].arrival <= now:
                self.waiting.append(pending[k])
                k += 1
            self._admit(now)
            if not self.active:
                if k < len(pending):  # idle until the next arrival
                    time.sleep(max(0.0, min(0.01, pending[k].arrival - now)))
                continue
            self._feed(now)
            t_step = self.clock()
            out = self.engine.step()
            steps += 1
            self.step_times.append(self.clock() - t_step)
            self.feed_times.append(t_step - t0 - now)
            t_done = self.clock() - t0
            finished = []
            for c in out:
                run = self.active[c.session]
                seq = c.tag
                lat = t_done - run.submit_time.pop(seq)
                fd = FrameDone(run.req.rid, run.order[seq], seq, c.status, lat, c.frame)
                self.frames.append(fd)
                if self.on_frame:
                    self.on_frame(fd)
                run.completed += 1
                if not run.first_done:
                    run.first_done = True
                    self.first_frame.append(t_done - run.req.arrival)
                if run.completed == len(run.order):
                    finished.append(run)
            for run in finished:
                self.engine.close_session(run.sid)
                self.used -= run.nbytes
                del self.active[run.sid]
                self.done.append(run.req.rid)
        elapsed = self.clock() - t0
        def pct(v, q):
            v = sorted(v)
            return v[min(len(v) - 1, int(q * len(v)))] if v else 0.0

        lats = [f.latency_s for f in self.frames]
        return {"frames": len(self.frames), "elapsed_s": elapsed, "frames_per_s": len(self.frames) / elapsed,
                "p50_frame_ms": pct(lats, 0.5) * 1e3, "p99_frame_ms": pct(lats, 0.99) * 1e3,
                "admission_wait_p50_ms": pct(self.waits, 0.5) * 1e3,
                "admission_wait_p99_ms": pct(self.waits, 0.99) * 1e3,
                "first_frame_p50_ms": pct(self.first_frame, 0.5) * 1e3,
                "first_frame_p99_ms": pct(self.first_frame, 0.99) * 1e3, "steps": steps,
                "step_p50_ms": pct(self.step_times, 0.5) * 1e3, "step_p99_ms": pct(self.step_times, 0.99) * 1e3,
                "admit_feed_p50_ms": pct(self.feed_times, 0.5) * 1e3,
                "requests": len(pending), "max_active_sessions": self.max_active,
                "peak_session_bytes": self.peak_bytes, "errors": sum(1 for f in self.frames if f.status)}


def poisson_requests(streams, n: int | None, rate_per_s: float, seed: int, seek_frac: float = 0.2,
                     loop_frac: float = 0.1, min_len: int = 30, max_len: int = 600, until_s: float | None = None):
    """A seeded request mix: Poisson arrivals at `rate_per_s` (n requests, or
    every arrival before `until_s` seconds); streams drawn uniformly; lengths
    uniform in [min_len, max_len] (clamped to the stream for plays, beyond it
    for loop restarts); `seek_frac` of the requests start at a random key
    frame, `loop_frac` play past the end and restart."""
    import numpy as np

    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    rid = -1
    while n is None or rid + 1 < n:
        rid += 1
        t += float(rng.exponential(1.0 / rate_per_s))
        if until_s is not None and t >= until_s:
            break
        st = streams[int(rng.integers(0, len(streams)))]
        nf = len(st.packets)
        kind = rng.random()
        length = int(rng.integers(min_len, max_len + 1))
        seek_at, seek_key = None, 0
        if kind < seek_frac and len(st.keys) > 1:
            # play a while, then seek to a random key frame and play on
            seek_key = int(st.keys[int(rng.integers(1, len(st.keys)))])
            seek_at = int(rng.integers(1, max(2, min(length, nf) // 2 + 1)))
            count = seek_at + min(length, nf - seek_key)
        elif kind < seek_frac + loop_frac:
            count = max(length, nf + 1)  # past the end: restart at frame 0
        else:
            count = min(length, nf)
        out.append(Request(rid, st, 0, count, t, seek_at, seek_key))
    return out


__all__ = ["Stream", "Request", "FrameDone", "Server", "poisson_requests"]
