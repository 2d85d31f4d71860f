"""Seeded synthetic clip generators.

The reference ships no EBIV streams; every EBIV test synthesises content in
process (ebiv/harness_test.go:105-171).  This module restates its two
generators (synthFrame, synthPan) and defines the workloads the benchmark and
the parity tests use:

* ``rects``     — smooth seeded gradient background, four textured rectangles
                  (uniform-random RGB noise textures, 16-48 px) moving at seeded
                  integer or half-pel velocities, Gaussian noise sigma 2.
* ``parallax``  — three seeded Perlin-noise layers scrolling at 0.5 / 1 / 2 px per
                  frame, a 2-frame brightness flash every 24 frames, Gaussian
                  noise sigma 1.5.
* ``noisepan``  — mid-grey base plus uniform noise (sigma 20) with a random
                  per-frame global translation in [-16, 16] px.

Seeds: splitmix64 of (master, workload, stream index); numpy's PCG64 draws the
values; Gaussians come from numpy's normal.  RGB is converted with the JFIF
fixed-point formula of Go's color.RGBToYCbCr, and chroma is the rounded mean of
each 2x2 block (edge-replicated for odd sizes).
"""
from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1
MASTER_SEED = 0xEB1F_2026_0001


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def stream_seed(workload: str, index: int, master: int = MASTER_SEED) -> int:
    h = master
    for ch in workload.encode():
        h = splitmix64(h ^ ch)
    return splitmix64(h ^ (index & MASK64))


# ---- harness_test.go:105-171 ---------------------------------------------------------
def synth_frame(w: int, h: int, seed: int):
    """synthFrame: gradient, hard diagonal edge, 4x4 checker, independent chroma."""
    y, x = np.mgrid[0:h, 0:w]
    v = (x * 3 + y * 5 + seed * 7) & 0xFF
    v = np.where(x + y > (w + h) // 2, 255 - v, v)
    v = np.where(((x // 4 + y // 4) & 1) == 0, v ^ 0x33, v)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    cy, cx = np.mgrid[0:ch, 0:cw]
    cb = (cx * 11 + cy * 2 + seed * 3) & 0xFF
    cr = (255 - cx * 2 - cy * 13 - seed * 5) & 0xFF
    return v.astype(np.uint8), cb.astype(np.uint8), cr.astype(np.uint8)


def synth_pan(w: int, h: int, seed: int):
    """synthPan: fixed texture translated by (2i, 2i) luma pixels."""
    off = 2 * seed
    y, x = np.mgrid[0:h, 0:w]
    X, Y = x + off, y + off
    v = (X * 3 + Y * 2) & 0xFF
    v = np.where(((X // 6 + Y // 6) & 1) == 0, v ^ 0x40, v)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    cy, cx = np.mgrid[0:ch, 0:cw]
    CX, CY = cx + seed, cy + seed
    cb = (CX * 5 + CY * 3) & 0xFF
    cr = (200 - CX * 2 - CY * 4) & 0xFF
    return v.astype(np.uint8), cb.astype(np.uint8), cr.astype(np.uint8)


# ---- colour conversion -------------------------------------------------------------------
def rgb_to_ycbcr420(rgb: np.ndarray):
    """color.RGBToYCbCr fixed point per pixel, then 2x2-mean chroma."""
    r = rgb[..., 0].astype(np.int32)
    g = rgb[..., 1].astype(np.int32)
    b = rgb[..., 2].astype(np.int32)
    yy = (19595 * r + 38470 * g + 7471 * b + (1 << 15)) >> 16

    def fix(v):  # uint32(v)&0xff000000 == 0 ? v>>16 : ^(v>>31)
        return np.where(v < 0, 0, np.where(v >= (1 << 24), 255, v >> 16))

    cb = fix(-11056 * r - 21712 * g + 32768 * b + (257 << 15))
    cr = fix(32768 * r - 27440 * g - 5328 * b + (257 << 15))
    h, w = yy.shape
    ph, pw = h + (h & 1), w + (w & 1)

    def sub(c):
        c = np.pad(c, ((0, ph - h), (0, pw - w)), mode="edge")
        s = c[0::2, 0::2] + c[1::2, 0::2] + c[0::2, 1::2] + c[1::2, 1::2]
        return ((s + 2) >> 2).astype(np.uint8)

    return yy.astype(np.uint8), sub(cb), sub(cr)


def _to_u8(x: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(x), 0, 255, out=x if x.dtype == np.float32 else None).astype(np.uint8)


# ---- gradient noise ------------------------------------------------------------------------
def perlin(h: int, w: int, rng: np.random.Generator, cell: int = 64, octaves: int = 4) -> np.ndarray:
    """Fractal 2-D gradient noise in roughly [-1, 1], shape (h, w)."""
    out = np.zeros((h, w), np.float32)
    amp, total = 1.0, 0.0
    for o in range(octaves):
        c = max(2, cell >> o)
        gh, gw = h // c + 2, w // c + 2
        ang = rng.uniform(0, 2 * np.pi, (gh, gw))
        gx, gy = np.cos(ang), np.sin(ang)
        ys = np.arange(h, dtype=np.float32) / c
        xs = np.arange(w, dtype=np.float32) / c
        y0 = np.floor(ys).astype(np.int32)[:, None]
        x0 = np.floor(xs).astype(np.int32)[None, :]
        fy = (ys[:, None] - y0).astype(np.float32)
        fx = (xs[None, :] - x0).astype(np.float32)
        gx32, gy32 = gx.astype(np.float32), gy.astype(np.float32)

        def dot(iy, ix, dy, dx):
            return gx32[iy, ix] * dx + gy32[iy, ix] * dy

        n00 = dot(y0, x0, fy, fx)
        n10 = dot(y0, x0 + 1, fy, fx - 1)
        n01 = dot(y0 + 1, x0, fy - 1, fx)
        n11 = dot(y0 + 1, x0 + 1, fy - 1, fx - 1)
        sx = fx * fx * fx * (fx * (fx * 6 - 15) + 10)
        sy = fy * fy * fy * (fy * (fy * 6 - 15) + 10)
        nx0 = n00 + sx * (n10 - n00)
        nx1 = n01 + sx * (n11 - n01)
        out += np.float32(amp) * (nx0 + sy * (nx1 - nx0))
        total += amp
        amp *= 0.5
    return out * np.float32(1.4 / total)


# ---- workloads ------------------------------------------------------------------------------
class RectsClip:
    """Gradient background + 4 noise-textured rectangles + sigma-2 noise."""

    def __init__(self, w: int, h: int, seed: int, noise_sigma: float = 2.0):
        self.w, self.h = w, h
        self.rng = np.random.default_rng(seed)
        rng = self.rng
        corners = rng.uniform(30, 225, (2, 2, 3))
        ty = np.linspace(0, 1, h)[:, None, None]
        tx = np.linspace(0, 1, w)[None, :, None]
        top = corners[0, 0] * (1 - tx) + corners[0, 1] * tx
        bot = corners[1, 0] * (1 - tx) + corners[1, 1] * tx
        self.bg = top * (1 - ty) + bot * ty
        self.rects = []
        for _ in range(4):
            rw, rh = int(rng.integers(16, 49)), int(rng.integers(16, 49))
            tex = rng.integers(0, 256, (rh, rw, 3)).astype(np.float64)
            pos = (int(rng.integers(0, 2 * w)), int(rng.integers(0, 2 * h)))  # half-pel units
            vel = (int(rng.integers(-4, 5)), int(rng.integers(-4, 5)))        # half-pel / frame
            self.rects.append((tex, pos, vel))
        self.sigma = noise_sigma

    def frame(self, t: int):
        w, h = self.w, self.h
        img = self.bg.copy()
        for tex, (px, py), (vx, vy) in self.rects:
            hx = (px + vx * t) % (2 * w)
            hy = (py + vy * t) % (2 * h)
            x0, y0, fx, fy = hx >> 1, hy >> 1, hx & 1, hy & 1
            rh, rw = tex.shape[:2]
            a = np.zeros((rh + 1, rw + 1), np.float64)
            p = np.zeros((rh + 1, rw + 1, 3), np.float64)
            for dy, wy in ((0, 1 - 0.5 * fy), (1, 0.5 * fy)):
                for dx, wx in ((0, 1 - 0.5 * fx), (1, 0.5 * fx)):
                    if wx * wy == 0:
                        continue
                    a[dy:dy + rh, dx:dx + rw] += wx * wy
                    p[dy:dy + rh, dx:dx + rw] += wx * wy * tex
            ys = (np.arange(rh + 1) + y0) % h
            xs = (np.arange(rw + 1) + x0) % w
            sub = img[np.ix_(ys, xs)]
            img[np.ix_(ys, xs)] = sub * (1 - a[..., None]) + p
        img = img + self.rng.normal(0, self.sigma, img.shape)
        return rgb_to_ycbcr420(_to_u8(img))


class ParallaxClip:
    """Three Perlin layers scrolling 0.5 / 1 / 2 px per frame, flash every 24."""

    SPEEDS = (0.5, 1.0, 2.0)

    def __init__(self, w: int, h: int, seed: int, frames: int, noise_sigma: float = 1.5):
        self.w, self.h = w, h
        self.rng = np.random.default_rng(seed)
        self.layers = []
        for i, sp in enumerate(self.SPEEDS):
            tw = w + int(np.ceil(sp * frames)) + 2
            base = self.rng.uniform(40, 215, 3)
            tint = self.rng.uniform(20, 60, 3)
            n = perlin(h, tw, self.rng, cell=96 >> i)
            tex = (base[None, None, :] + n[..., None] * (tint * 1.5)[None, None, :]).astype(np.float32)
            mask = None
            if i > 0:
                m = perlin(h, tw, self.rng, cell=128)
                mask = (m > 0.15 * i).astype(np.float32)
            self.layers.append((sp, tex, mask))
        self.sigma = noise_sigma

    def frame(self, t: int):
        w = self.w
        img = None
        for sp, tex, mask in self.layers:
            off = sp * t
            i0 = int(np.floor(off))
            f = np.float32(off - i0)
            layer = tex[:, i0:i0 + w] if f == 0 else tex[:, i0:i0 + w] * (1 - f) + tex[:, i0 + 1:i0 + 1 + w] * f
            if mask is None:
                img = layer.copy()
            else:
                m = mask[:, i0:i0 + w] if f == 0 else mask[:, i0:i0 + w] * (1 - f) + mask[:, i0 + 1:i0 + 1 + w] * f
                img += (layer - img) * m[..., None]
        if t % 24 in (12, 13):
            img += np.float32(48.0)
        img += self.rng.standard_normal(img.shape, dtype=np.float32) * np.float32(self.sigma)
        return rgb_to_ycbcr420(_to_u8(img))


class NoisePanClip:
    """Mid-grey + uniform noise (sigma 20) under random global motion in +-16 px."""

    MARGIN = 256

    def __init__(self, w: int, h: int, seed: int, sigma: float = 20.0):
        self.w, self.h = w, h
        rng = np.random.default_rng(seed)
        half = sigma * np.sqrt(3.0)
        self.tex = _to_u8(128 + rng.uniform(-half, half, (h + self.MARGIN, w + self.MARGIN, 3)))
        self.rng = rng
        self.pos = [self.MARGIN // 2, self.MARGIN // 2]

    def frame(self, t: int):
        if t > 0:
            for k in range(2):
                self.pos[k] = int(np.clip(self.pos[k] + self.rng.integers(-16, 17), 0, self.MARGIN))
        y0, x0 = self.pos
        return rgb_to_ycbcr420(self.tex[y0:y0 + self.h, x0:x0 + self.w])


def generate(workload: str, w: int, h: int, frames: int, seed: int):
    """Yields `frames` packed (Y, Cb, Cr) frames of a named workload."""
    if workload == "rects":
        clip = RectsClip(w, h, seed)
    elif workload == "parallax":
        clip = ParallaxClip(w, h, seed, frames)
    elif workload == "noisepan":
        clip = NoisePanClip(w, h, seed)
    elif workload == "synthframe":
        for i in range(frames):
            yield synth_frame(w, h, i)
        return
    elif workload == "synthpan":
        for i in range(frames):
            yield synth_pan(w, h, i)
        return
    else:
        raise ValueError(workload)
    for t in range(frames):
        yield clip.frame(t)
