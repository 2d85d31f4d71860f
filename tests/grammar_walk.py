"""Independent grammar walk over the encoder's token streams (test
infrastructure).

A restatement of the reference's bit-audit walker (ebiv/bitaudit_test.go:
306-441: auditWalker.next/nextAs/classValue/walkStream/walkIntraMB/
walkInterMB/block) and its pricing (makeCostTable, entropyBits, :180-224).
It re-parses a tile's (ctx, sym) tokens with the decoder's grammar
(decode_intra.go:121-167, decode_inter.go:8-118, residual.go:77-126,
mv.go:96-128) without touching pixels or rANS, and goes further than the
reference's walker in one respect: every token's context is checked against
the one the grammar expects at that point (model.go:124-190: tokenCtx,
bandOf, classForMag), so an encoder that emits the right symbols under a
wrong context fails here, not only by desynchronising a decoder written by
the same hand.

The walk yields per macroblock what the grammar fixes (mode, transform size,
intra modes, reference, partition, MV deltas, coded-block pattern, number of
nonzero levels), the bits of every token under the frame's final tables and
the token count, which is the number of rANS decodes a decoder makes.
"""
from __future__ import annotations

import numpy as np

# model.go:20-149
MB_SIZE, CHROMA_MB = 16, 8
LUMA_TX_SIZES = (4, 8, 16)
NUM_INTRA_MODES = 10
MB_SKIP, MB_INTER, MB_INTRA = 0, 1, 2
NUM_REFS = 2
CBP_LUMA, CBP_CB, CBP_CR = 1, 2, 4
PART_COUNT = (1, 2, 2, 4)  # partRects: whole, 16x8, 8x16, 8x8
T_ZERO, T_ESCAPE, T_EOB = 0, 5, 6
NUM_MV_CLASSES = NUM_ESC_CLASSES = 16
(CTX_TX, CTX_LUMA_MODE, CTX_CHROMA_MODE, CTX_MB_MODE, CTX_REF, CTX_PART, CTX_CBP, CTX_SIGN, CTX_ESC,
 CTX_MVX, CTX_MVY, CTX_BYPASS, CTX_TOKEN_BASE) = range(13)
NUM_BANDS, NUM_PREV_CLASS = 5, 3
LUMA_TOKEN_CTX = 3 * NUM_BANDS * NUM_PREV_CLASS
NUM_CONTEXTS = CTX_TOKEN_BASE + LUMA_TOKEN_CTX + NUM_BANDS * NUM_PREV_CLASS
SDH_MIN_SPAN = 4  # residual.go:25
RANS_M = 4096     # rans.go: probability scale

GROUP_OF = {CTX_MB_MODE: "mbmode", CTX_REF: "ref", CTX_PART: "part", CTX_TX: "txsize", CTX_CBP: "cbp",
            CTX_LUMA_MODE: "mode.luma", CTX_CHROMA_MODE: "mode.chroma", CTX_SIGN: "sign", CTX_ESC: "escape",
            CTX_MVX: "mv", CTX_MVY: "mv"}


def token_ctx(plane: int, tx: int, band: int, prev_class: int) -> int:
    """tokenCtx, model.go:155-161 (plane 0 luma, 1 chroma)."""
    if plane == 0:
        return CTX_TOKEN_BASE + (tx * NUM_BANDS + band) * NUM_PREV_CLASS + prev_class
    return CTX_TOKEN_BASE + LUMA_TOKEN_CTX + band * NUM_PREV_CLASS + prev_class


def band_of(pos: int) -> int:
    """bandOf, model.go:165-178."""
    return 0 if pos == 0 else 1 if pos <= 2 else 2 if pos <= 8 else 3 if pos <= 20 else 4


def zigzag_decode(v: int) -> int:
    """zigzagDecode (motioncomp.go:410-412): 0, -1, 1, -2, 2, ... from 0, 1, 2, 3, 4, ..."""
    return (v >> 1) ^ -(v & 1)


class GrammarError(AssertionError):
    pass


def cost_table(freqs: np.ndarray):
    """makeCostTable (bitaudit_test.go:180-202): bits per (ctx, sym) under the
    frame's tables; the bypass context costs one bit per symbol; a symbol of
    zero frequency has no cost (it cannot have been coded: checked apart)."""
    f = freqs.astype(np.float64)
    with np.errstate(divide="ignore"):
        c = np.where(f > 0, np.log2(RANS_M / np.maximum(f, 1)), 0.0)
    c[CTX_BYPASS, :] = 0.0
    c[CTX_BYPASS, :2] = 1.0
    return c


def entropy_bits(counts: np.ndarray, freqs: np.ndarray) -> float:
    """entropyBits (bitaudit_test.go:204-224): sum over (ctx, sym) of count x
    cost, the bypass context at one bit per symbol."""
    bits = float(counts[CTX_BYPASS].sum())
    f = freqs.astype(np.float64)
    for c in range(NUM_CONTEXTS):
        if c == CTX_BYPASS:
            continue
        m = (counts[c] > 0) & (f[c] > 0)
        bits += float((counts[c][m] * np.log2(RANS_M / f[c][m])).sum())
    return bits


class TileWalk:
    """auditWalker over one tile's tokens (bitaudit_test.go:306-441)."""

    def __init__(self, toks: np.ndarray, cost: np.ndarray):
        self.ctx = toks[:, 0].astype(np.int64).tolist()
        self.sym = toks[:, 1].astype(np.int64).tolist()
        self.cost = cost
        self.pos = 0
        self.bits = 0.0
        self.groups: dict = {}
        self.mbs: list = []

    # -- token access -------------------------------------------------------
    def _take(self, ctx: int, group: str | None) -> int:
        if self.pos >= len(self.ctx):
            raise GrammarError(f"token {self.pos}: stream ends inside a macroblock (wanted ctx {ctx})")
        c, s = self.ctx[self.pos], self.sym[self.pos]
        if c != ctx:
            raise GrammarError(f"token {self.pos}: ctx {c}, the grammar expects {ctx}")
        self.pos += 1
        b = float(self.cost[c, s]) if s < self.cost.shape[1] else 0.0
        self.bits += b
        if group is None:  # record(): the context's own group
            group = GROUP_OF.get(c) or ("coef.luma" if c < CTX_TOKEN_BASE + LUMA_TOKEN_CTX else "coef.chroma")
        self.groups[group] = self.groups.get(group, 0.0) + b
        return s

    def next(self, ctx: int) -> int:
        return self._take(ctx, None)

    def next_as(self, ctx: int, group: str) -> int:
        """nextAs: bypass suffix bits carry the group the grammar names."""
        return self._take(ctx, group)

    def class_value(self, class_ctx: int, max_class: int, group: str) -> int:
        """classValue / getClassValue (mv.go:96-110): class token, then
        class-1 raw suffix bits MSB first through the bypass context."""
        cls = self.next(class_ctx)
        if cls >= max_class:
            raise GrammarError(f"token {self.pos - 1}: class {cls} out of range")
        if cls <= 1:
            return cls
        v = 1 << (cls - 1)
        for i in range(cls - 1):
            v |= self.next_as(CTX_BYPASS, group) << (cls - 2 - i)
        return v

    # -- grammar ------------------------------------------------------------
    def block(self, plane: int, tx: int, n: int) -> int:
        """One residual block (residual.go:79-126); returns its nonzero count.
        Signs are inline after the first nonzero; the first sign comes at the
        end when the nonzero span is below SDH_MIN_SPAN."""
        first = last = -1
        prev_class = 0
        nz = 0
        for i in range(n * n):
            t = self.next(token_ctx(plane, tx, band_of(i), prev_class))
            if t == T_EOB:
                break
            if t == T_ZERO:
                prev_class = 0
                continue
            if t == T_ESCAPE:
                self.class_value(CTX_ESC, NUM_ESC_CLASSES, "escape")
                prev_class = 2
            elif T_ZERO < t < T_ESCAPE:
                prev_class = 1 if t == 1 else 2
            else:
                raise GrammarError(f"token {self.pos - 1}: coefficient symbol {t}")
            nz += 1
            if first < 0:
                first = i
            else:
                self.next(CTX_SIGN)
            last = i
        if first >= 0 and last - first < SDH_MIN_SPAN:
            self.next(CTX_SIGN)
        return nz

    def intra_mb(self) -> dict:
        """walkIntraMB: transform size, per luma block its mode and levels,
        the chroma mode, then the Cb and Cr blocks."""
        tx = self.next(CTX_TX)
        if tx >= len(LUMA_TX_SIZES):
            raise GrammarError(f"token {self.pos - 1}: tx {tx}")
        n = LUMA_TX_SIZES[tx]
        lmodes, nz = [], 0
        for _ in range((MB_SIZE // n) ** 2):
            m = self.next(CTX_LUMA_MODE)
            if m >= NUM_INTRA_MODES:
                raise GrammarError(f"token {self.pos - 1}: luma mode {m}")
            lmodes.append(m)
            nz += self.block(0, tx, n)
        cmode = self.next(CTX_CHROMA_MODE)
        if cmode >= NUM_INTRA_MODES:
            raise GrammarError(f"token {self.pos - 1}: chroma mode {cmode}")
        nz += self.block(1, 0, CHROMA_MB)
        nz += self.block(1, 0, CHROMA_MB)
        return {"kind": MB_INTRA, "tx": tx, "lmode": lmodes, "cmode": cmode, "ncoef": nz}

    def inter_mb(self) -> dict:
        """walkInterMB: mode; skip ends it; intra as a key-frame macroblock;
        inter: reference, partition, per partition the x and y MV deltas, the
        coded-block pattern, then the coded components' blocks."""
        mode = self.next(CTX_MB_MODE)
        if mode == MB_SKIP:
            return {"kind": MB_SKIP, "ncoef": 0}
        if mode == MB_INTRA:
            return self.intra_mb()
        if mode != MB_INTER:
            raise GrammarError(f"token {self.pos - 1}: macroblock mode {mode}")
        ref = self.next(CTX_REF)
        part = self.next(CTX_PART)
        if ref >= NUM_REFS or part >= len(PART_COUNT):
            raise GrammarError(f"token {self.pos - 1}: ref {ref} part {part}")
        mvd = []
        for _ in range(PART_COUNT[part]):
            dx = zigzag_decode(self.class_value(CTX_MVX, NUM_MV_CLASSES, "mv"))
            dy = zigzag_decode(self.class_value(CTX_MVY, NUM_MV_CLASSES, "mv"))
            mvd.append((dx, dy))
        cbp = self.next(CTX_CBP)
        if cbp >= 8:
            raise GrammarError(f"token {self.pos - 1}: cbp {cbp}")
        tx, nz = -1, 0
        if cbp & CBP_LUMA:
            tx = self.next(CTX_TX)
            if tx >= len(LUMA_TX_SIZES):
                raise GrammarError(f"token {self.pos - 1}: tx {tx}")
            n = LUMA_TX_SIZES[tx]
            for _ in range((MB_SIZE // n) ** 2):
                nz += self.block(0, tx, n)
        if cbp & CBP_CB:
            nz += self.block(1, 0, CHROMA_MB)
        if cbp & CBP_CR:
            nz += self.block(1, 0, CHROMA_MB)
        return {"kind": MB_INTER, "ref": ref, "part": part, "mvd": mvd, "cbp": cbp, "tx": tx, "ncoef": nz}

    def walk(self, inter: bool, n_mbs: int) -> list:
        """walkStream for a tile of n_mbs macroblocks: the stream must hold
        exactly that many macroblocks and nothing after them."""
        for _ in range(n_mbs):
            self.mbs.append(self.inter_mb() if inter else self.intra_mb())
        if self.pos != len(self.ctx):
            raise GrammarError(f"walker consumed {self.pos} of {len(self.ctx)} tokens: grammar mismatch")
        return self.mbs


def tile_bounds(mbc: int, mbr: int, cols: int, rows: int):
    """newTileGrid (frame.go:87-111): edges at mbc*i/cols, mbr*i/rows; tiles
    in raster order as (x0, x1, y0, y1) in macroblocks."""
    ce = [mbc * i // cols for i in range(cols + 1)]
    re_ = [mbr * i // rows for i in range(rows + 1)]
    return [(ce[c], ce[c + 1], re_[r], re_[r + 1]) for r in range(rows) for c in range(cols)]


def walk_payload(p) -> dict:
    """Walks every tile of an encoder Payload. Returns the per-macroblock
    records in frame raster order, the token count, walked bits, the
    entropy of the counts, per-tile bits and the group tally."""
    cost = cost_table(p.freqs)
    mbc, mbr = p.mbs
    recs: list = [None] * (mbc * mbr)
    counts = np.zeros((NUM_CONTEXTS, 16), np.int64)
    tokens, bits, tile_bits, groups = 0, 0.0, [], {}
    for t, (toks, (x0, x1, y0, y1)) in enumerate(zip(p.tiles, tile_bounds(mbc, mbr, *p.grid))):
        w = TileWalk(toks, cost)
        try:
            mbs = w.walk(p.inter, (x1 - x0) * (y1 - y0))
        except GrammarError as e:
            raise GrammarError(f"packet {p.packet} tile {t}: {e}") from None
        k = 0
        for y in range(y0, y1):
            for x in range(x0, x1):
                recs[y * mbc + x] = mbs[k]
                k += 1
        np.add.at(counts, (toks[:, 0].astype(np.int64), toks[:, 1].astype(np.int64)), 1)
        tokens += len(toks)
        bits += w.bits
        tile_bits.append(w.bits)
        for g, b in w.groups.items():
            groups[g] = groups.get(g, 0.0) + b
    return {"recs": recs, "tokens": tokens, "bits": bits, "entropy": entropy_bits(counts, p.freqs),
            "tile_bits": tile_bits, "groups": groups, "counts": counts}

