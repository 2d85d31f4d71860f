"""Golden stream fixtures (tests/golden/): the oracle and the container reader
against recorded per-frame plane CRCs, and the encoder's determinism."""
import glob
import json
import os
import zlib

import pytest

import oracle as O
from govid_amd.container import Reader

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(HERE, "*.json")))


def load(name):
    with open(os.path.join(HERE, name + ".json")) as fh:
        meta = json.load(fh)
    return meta, Reader.open(os.path.join(HERE, name + ".ebiv"))


def test_fixtures_present():
    assert {"cfg1_rects_320x180", "altref_parallax_64x48", "noise_q12_128x96_2x2", "odd_37x21_2x3",
            "raw_48x32"} <= set(NAMES)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_recorded_crcs(name):
    meta, rd = load(name)
    assert (rd.header.width, rd.header.height) == (meta["width"], meta["height"])
    pkts = rd.packets()
    assert len(pkts) == meta["frames"]
    dec = O.OracleCodec()
    for i, p in enumerate(pkts):
        got = [zlib.crc32(a.tobytes()) for a in dec.decode(p)]
        assert got == meta["plane_crc32"][i], f"{name} frame {i}"


@pytest.mark.parametrize("name", ["altref_parallax_64x48", "odd_37x21_2x3", "raw_48x32"])
def test_encoder_reproduces_fixture_bytes(name):
    """The seeded generator and the encoder restatement are deterministic: the
    committed stream is rebuilt byte for byte."""
    import sys

    sys.path.insert(0, HERE)
    import make_fixtures as M
    from govid_amd import clips, encoder

    meta, _ = load(name)
    spec = M.FIXTURES[name]
    frames = list(clips.generate(spec["workload"], spec["w"], spec["h"], spec["frames"], meta["seed"]))
    clip = encoder.encode_frames(frames, spec["w"], spec["h"], threads=4, **spec["enc"])
    assert zlib.crc32(clip.container) == meta["container_crc32"]
