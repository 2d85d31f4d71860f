"""The benchmark streams' expected CRCs against the oracle decoder.

bench_data/make_streams.py stores, beside every benchmark stream, the CRC-32
of each frame's packed Y, Cb and Cr as the encoder reconstructed them; the
benchmark checks the GPU's frames against those files.  Here the oracle (the
C restatement of the Go decoder) decodes the streams and must give the same
CRCs: the format's contract that a decoder's output equals the encoder's
reconstruction (ebiv/coded_test.go:107-124), at benchmark scale.  Every frame
of the committed streams is checked; of the generated sets (bench_data/gen/,
present only where they were generated), the first frames of each stream."""
import glob
import json
import os
import sys
import zlib

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

from govid_amd.container import Reader  # noqa: E402

COMMITTED = sorted(glob.glob(os.path.join(ROOT, "bench_data", "*.ebiv")))
GENERATED = sorted(glob.glob(os.path.join(ROOT, "bench_data", "gen", "*.ebiv")))


def _check(path, limit=None):
    rd = Reader.open(path)
    with open(path[:-len(".ebiv")] + ".crc.json") as fh:
        exp = json.load(fh)
    assert exp["stream_crc32"] == zlib.crc32(rd.data)
    assert (exp["width"], exp["height"]) == (rd.header.width, rd.header.height)
    pk = rd.packets()
    assert len(exp["crc32"]) == len(pk)
    dec = O.OracleCodec(threads=8)
    for i, p in enumerate(pk[:limit]):
        got = [zlib.crc32(pl.tobytes()) for pl in dec.decode(p)]
        assert got == exp["crc32"][i], f"{os.path.basename(path)} frame {i}"


@pytest.mark.parametrize("path", COMMITTED, ids=os.path.basename)
def test_committed_stream_crcs_match_the_oracle(path):
    _check(path)


@pytest.mark.skipif(not GENERATED, reason="bench_data/gen/ not generated here")
def test_generated_stream_crcs_match_the_oracle():
    for path in GENERATED:
        _check(path, limit=20)
