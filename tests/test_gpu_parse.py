"""Entropy-kernel parity: the GPU's macroblock records and coefficient lists
against the oracle's per-macroblock trace of the same frame."""
import numpy as np
import pytest

import oracle as O
from helpers import clip_kw

pytestmark = pytest.mark.gpu

PART_COUNT = [1, 2, 2, 4]
ZIGZAG = {n: O.zigzag(n) for n in (4, 8, 16)}


def to_position(e, tx):
    """A GPU coefficient entry with its scan index replaced by the raster
    position (oracle zigzag of the block's size; chroma blocks are 8x8)."""
    si, blk = e & 255, (e >> 8) & 31
    n = 8 if blk >= 16 else 4 << tx
    return (e & ~255) | ZIGZAG[n][si]


def compare_frame(gpu, ref, label=""):
    recs, coefs, status = gpu.debug_dump(0)
    tr = ref.trace()
    bad = []
    for i, t in enumerate(tr):
        if t["kind"] < 0:
            continue
        r = recs[i]
        errs = []
        if int(r["kind"]) != t["kind"]:
            errs.append(f"kind {r['kind']} != {t['kind']}")
        else:
            k = t["kind"]
            if k == 2:
                n = 4 << t["tx"]
                nb = (16 // n) ** 2
                if int(r["tx"]) != t["tx"]:
                    errs.append(f"tx {r['tx']} != {t['tx']}")
                lm = [(int(r["lmode"][b // 8]) >> (4 * (b % 8))) & 15 for b in range(nb)]
                if lm != t["lmode"][:nb]:
                    errs.append(f"lmodes {lm} != {t['lmode'][:nb]}")
                if int(r["cmode"]) != t["cmode"]:
                    errs.append(f"cmode {r['cmode']} != {t['cmode']}")
            elif k == 1:
                np_ = PART_COUNT[t["part"]]
                gm = r["mv"][:np_].flatten().tolist()
                if int(r["part"]) != t["part"] or int(r["refsel"]) != t["refsel"] or gm != t["mv"][:2 * np_]:
                    errs.append(f"part/ref/mv {r['part']},{r['refsel']},{gm} != {t['part']},{t['refsel']},{t['mv'][:2*np_]}")
                if int(r["cbp"]) != t["cbp"]:
                    errs.append(f"cbp {r['cbp']} != {t['cbp']}")
            else:
                if r["mv"][0].tolist() != t["mv"][:2]:
                    errs.append(f"skip mv {r['mv'][0].tolist()} != {t['mv'][:2]}")
            off, cnt = int(r["coef_off"]), int(r["coef_count"])
            # the GPU stores scan indices, the oracle's trace raster positions
            g = sorted(to_position(e, int(r["tx"])) for e in coefs[off:off + cnt].tolist())
            if g != t["coef"]:
                gs, ts = set(g), set(t["coef"])
                dec = lambda e: (e & 255, (e >> 8) & 31, np.int32(np.uint32(e)) >> 13)
                errs.append(f"coefs: {len(g)} vs {len(t['coef'])}; gpu-only {[dec(e) for e in sorted(gs - ts)][:6]} "
                            f"oracle-only {[dec(e) for e in sorted(ts - gs)][:6]}")
        if errs:
            bad.append((i, errs))
    return bad


def run_clip(clip, frames=None, label=""):
    from govid_amd.codec import Codec

    gpu = Codec()
    ref = O.OracleCodec(threads=8)
    ref.enable_trace(True)
    W = clip.width if hasattr(clip, "width") else None
    for fi, pkt in enumerate(clip.packets[: frames or len(clip.packets)]):
        gpu.decode(pkt)
        ref.decode(pkt)
        bad = compare_frame(gpu, ref)
        if bad:
            mbc = (clip.width + 15) // 16
            lines = [f"MB ({i % mbc},{i // mbc}): {'; '.join(e)}" for i, e in bad[:8]]
            pytest.fail(f"{label} frame {fi}: {len(bad)} macroblocks differ\n" + "\n".join(lines))


@pytest.mark.parametrize("case", [
    dict(workload="synthframe", w=64, h=48, frames=3, qp=20, gop=1),
    dict(workload="rects", w=96, h=64, frames=6, qp=22, gop=6, tile_cols=2, tile_rows=2),
    dict(workload="parallax", w=256, h=128, frames=6, qp=22, gop=6, tile_cols=2, tile_rows=1),
    # single tiles wider than 64 macroblocks (MV rows in LDS), 65 and 72 wide
    dict(workload="rects", w=1040, h=64, frames=4, qp=22, gop=4, tile_cols=1, tile_rows=1),
    dict(workload="synthpan", w=1152, h=80, frames=5, qp=18, gop=5, tile_cols=1, tile_rows=1),
])
def test_parse_matches_oracle(case):
    c = dict(case)
    wl, w, h, n = c.pop("workload"), c.pop("w"), c.pop("h"), c.pop("frames")
    print(f"case {wl} {w}x{h} x{n} {c}", flush=True)
    run_clip(clip_kw(wl, w, h, n, threads=4, **c), label=wl)


def test_parse_bench_clip_first_frames():
    import os
    from govid_amd.container import Reader

    path = os.path.join(os.path.dirname(os.path.dirname(__file__)), "bench_data", "parallax_1080p.ebiv")
    rd = Reader.open(path)

    class C:
        packets = rd.packets()[:3]
        width = rd.header.width

    run_clip(C, label="parallax_1080p")


def test_generic_grammar_decoder_without_a_sign_table():
    """A frame with no sign table takes the entropy kernel's generic grammar
    decoder (parse_tile) rather than the hand-scheduled token loop, and the
    stats count those tiles.  The encoder leaves the table out when a frame
    codes no coefficient (flat content: key-frame modes, then skips); a coded
    key frame whose sign table was removed fails at its first sign symbol
    (unused context, rans.go:275-278) after the generic decoder has read its
    first blocks.  Records, planes and statuses match the oracle."""
    from govid_amd import codec as K
    from govid_amd import encoder
    from govid_amd.codec import Codec
    from helpers import assert_planes_equal, edit_tables

    w, h = 96, 64
    flat = [(np.full((h, w), 128, np.uint8), np.full((h // 2, w // 2), 128, np.uint8),
             np.full((h // 2, w // 2), 128, np.uint8)) for _ in range(4)]
    clip = encoder.encode_frames(flat, w, h, qp=22, gop=2, tile_cols=2, tile_rows=2)
    print("flat clip", flush=True)
    run_clip(clip, label="flat")
    gpu, ref = Codec(ring=3), O.OracleCodec(ring=3)
    for i, p in enumerate(clip.packets):
        assert_planes_equal(gpu.decode(p).planes(), ref.decode(p), f"flat frame {i}")
    st = gpu.stats()
    assert st["tiles"] == 4 * len(clip.packets) and st["tiles_generic"] == st["tiles"], st

    coded = list(clip_kw("rects", w, h, 4, qp=22, gop=2, tile_cols=2, tile_rows=2).packets)
    CTX_SIGN = 7  # the sign context among the table block's 72
    coded[2] = edit_tables(coded[2], lambda e: [(0, []) if c == CTX_SIGN else ev for c, ev in enumerate(e)])
    gpu2, ref2 = Codec(ring=3), O.OracleCodec(ring=3)
    for i, p in enumerate(coded):
        print(f"  coded frame {i}", flush=True)
        try:
            want = ref2.decode(p)
        except O.DecodeError as e:
            with pytest.raises(K.EbivError) as got:
                gpu2.decode(p)
            status = {K.ErrCorrupt: 1, K.ErrUnsupported: 2, K.ErrNoKeyframe: 3, K.ErrDimensions: 4}[type(got.value)]
            assert status == e.status, (i, got.value, e)
            continue
        assert_planes_equal(gpu2.decode(p).planes(), want, f"coded frame {i}")
    st2 = gpu2.stats()
    # (the frame after it fails on the host: its tables delta-code the removed one)
    assert st2["tiles_generic"] == 4 and st2["tiles"] == 4 * 3, st2


def test_parse_matches_the_grammar_walk():
    """The entropy kernel against the independent grammar walk of the
    encoder's own token streams (ebiv/bitaudit_test.go:306-441), not against
    the oracle: per tile the GPU's rANS decode count equals the walker's
    token count, per packet (override payloads included) the codec's symbol
    count does, and per macroblock the records carry the walker's mode,
    transform size, intra modes, reference, partition, coded-block pattern
    and nonzero count."""
    from govid_amd.codec import Codec
    from grammar_walk import MB_INTER, MB_INTRA, walk_payload
    from test_grammar_walk import CASES, _by_packet, _clip

    for case in CASES:
        print(f"case {case}", flush=True)
        clip = _clip(case)
        by_packet = _by_packet(clip)
        gpu = Codec()
        for i, pkt in enumerate(clip.packets):
            before = gpu.stats()["symbols"]
            gpu.decode(pkt)
            walks = [walk_payload(p) for p in by_packet[i]]
            assert gpu.stats()["symbols"] - before == sum(w["tokens"] for w in walks), f"packet {i}"
            recs, _, status = gpu.debug_dump(0)
            tiles = by_packet[i][0].tiles
            assert [int(s["nsym"]) for s in status[:len(tiles)]] == [len(t) for t in tiles], f"packet {i}"
            assert all(int(s["err"]) == 0 for s in status[:len(tiles)])
            for m, r in enumerate(walks[0]["recs"]):
                g, where = recs[m], f"{case['workload']} packet {i} mb {m}"
                assert int(g["kind"]) == r["kind"] and int(g["coef_count"]) == r["ncoef"], where
                if r["kind"] == MB_INTRA:
                    nb = len(r["lmode"])
                    lm = [(int(g["lmode"][b // 8]) >> (4 * (b % 8))) & 15 for b in range(nb)]
                    assert (int(g["tx"]), lm, int(g["cmode"])) == (r["tx"], r["lmode"], r["cmode"]), where
                elif r["kind"] == MB_INTER:
                    assert (int(g["refsel"]), int(g["part"]), int(g["cbp"])) == (r["ref"], r["part"], r["cbp"]), where
        gpu.close()
