"""The grammar cross-check (ebiv/bitaudit_test.go:306-441): the encoder's
token streams walked by an independent restatement of the decoder's grammar,
priced against each frame's final tables, and compared with what the oracle
decoder parsed from the same packets.

Three hands meet here: the C++ encoder that wrote the tokens, the walker in
grammar_walk.py and the C oracle. A rule misread the same way by encoder and
oracle still has to get past the walker, whose context check is stricter than
the reference's own walker.
"""
import numpy as np
import pytest

import oracle as O
from grammar_walk import (CTX_BYPASS, CTX_SIGN, GrammarError, MB_INTER, MB_INTRA, TileWalk, cost_table, tile_bounds,
                          walk_payload)
from helpers import clip_kw, tile_spans

CASES = [
    dict(workload="synthframe", w=64, h=48, frames=3, qp=20, gop=1),
    dict(workload="rects", w=96, h=64, frames=6, qp=22, gop=6, tile_cols=2, tile_rows=2, fast=False),
    dict(workload="parallax", w=128, h=96, frames=9, qp=28, gop=9, altref=4, tile_cols=3, tile_rows=2),
    dict(workload="noisepan", w=96, h=64, frames=5, qp=8, gop=5),
    dict(workload="synthpan", w=37, h=21, frames=5, qp=22, gop=5),
]


def _clip(case):
    c = dict(case)
    wl, w, h, n = c.pop("workload"), c.pop("w"), c.pop("h"), c.pop("frames")
    return clip_kw(wl, w, h, n, threads=4, log_tokens=True, **c)


def _by_packet(clip):
    """payloads grouped per packet, in payload order within the packet"""
    out = {}
    for p in clip.payloads:
        out.setdefault(p.packet, []).append(p)
    return out


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['workload']}_{c['w']}x{c['h']}")
def test_walk_consumes_every_token_and_prices_to_the_entropy(case):
    clip = _clip(case)
    assert len(clip.payloads) >= len(clip.packets)
    for pkt_i, payloads in _by_packet(clip).items():
        pkt = clip.packets[pkt_i]
        for k, p in enumerate(payloads):
            w = walk_payload(p)
            assert all(r is not None for r in w["recs"]), "a macroblock no tile covers"
            # makeCostTable/entropyBits cross-check: walked bits = entropy of the counts
            assert w["bits"] == pytest.approx(w["entropy"], rel=1e-12, abs=1e-6)
            assert sum(w["groups"].values()) == pytest.approx(w["bits"], rel=1e-12, abs=1e-6)
            # every coded symbol has a nonzero frequency in its context's table
            coded = w["counts"] > 0
            coded[CTX_BYPASS] = False
            assert not (coded & (p.freqs == 0)).any(), f"packet {pkt_i}: a coded symbol has frequency 0"
            # rANS codes within a few bytes of the information content: the
            # tile's bytes hold its walked bits plus the four 32-bit states
            spans = tile_spans(pkt, k)
            assert len(spans) == len(p.tiles)
            for t, ((_, n), bits) in enumerate(zip(spans, w["tile_bits"])):
                assert 8 * n >= bits - 1e-6, f"packet {pkt_i} tile {t}: {n} B below {bits:.1f} bits"
                assert 8 * n <= bits * 1.002 + 8 * 20, f"packet {pkt_i} tile {t}: {n} B for {bits:.1f} bits"


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['workload']}_{c['w']}x{c['h']}")
def test_walk_matches_what_the_oracle_parsed(case):
    """Per macroblock: mode, transform size, intra modes, reference,
    partition, coded-block pattern and nonzero count agree with the oracle's
    trace of the main payload; per packet the walker's token count equals
    the oracle's rANS decode count."""
    clip = _clip(case)
    ref = O.OracleCodec(threads=4)
    ref.enable_trace(True)
    by_packet = _by_packet(clip)
    kinds = set()
    for i, pkt in enumerate(clip.packets):
        before = ref.symbols()
        ref.decode(pkt)
        walks = [walk_payload(p) for p in by_packet[i]]
        assert ref.symbols() - before == sum(w["tokens"] for w in walks), f"packet {i}: rANS decodes"
        tr = ref.trace()
        main = walks[0]["recs"]
        assert len(tr) == len(main)
        for m, (t, r) in enumerate(zip(tr, main)):
            where = f"packet {i} mb {m}"
            kinds.add(r["kind"])
            assert t["kind"] == r["kind"], where
            assert len(t["coef"]) == r["ncoef"], where
            if r["kind"] == MB_INTRA:
                nb = len(r["lmode"])
                assert (t["tx"], t["lmode"][:nb], t["cmode"]) == (r["tx"], r["lmode"], r["cmode"]), where
            elif r["kind"] == MB_INTER:
                assert (t["refsel"], t["part"], t["cbp"]) == (r["ref"], r["part"], r["cbp"]), where
    if case["gop"] > 1:
        assert MB_INTER in kinds, "an inter clip without a single inter macroblock tests little"


def test_walker_rejects_planted_mistakes():
    """The walker is not vacuous: a token under a wrong context, a dropped
    token and a trailing token each fail it."""
    clip = _clip(CASES[3])
    p = next(p for p in clip.payloads if p.inter and len(p.tiles[0]) > 50)
    cost = cost_table(p.freqs)
    x0, x1, y0, y1 = tile_bounds(*p.mbs, *p.grid)[0]
    n_mbs = (x1 - x0) * (y1 - y0)
    toks = p.tiles[0]
    TileWalk(toks, cost).walk(True, n_mbs)  # the real stream walks
    signs = np.nonzero(toks[:, 0] == CTX_SIGN)[0]
    assert len(signs), "no sign token to plant a mistake on"
    bad = toks.copy()
    bad[signs[0], 0] = CTX_BYPASS
    with pytest.raises(GrammarError):
        TileWalk(bad, cost).walk(True, n_mbs)
    with pytest.raises(GrammarError):
        TileWalk(np.delete(toks, len(toks) // 2, axis=0), cost).walk(True, n_mbs)
    with pytest.raises(GrammarError):
        TileWalk(np.concatenate([toks, toks[:1]]), cost).walk(True, n_mbs)
