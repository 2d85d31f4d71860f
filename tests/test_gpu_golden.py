"""The golden fixtures decoded on the GPU, plain and batched, against the
recorded per-frame plane CRCs."""
import glob
import json
import os
import zlib

import pytest

from govid_amd.container import Reader

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(HERE, "*.json")))


def load(name):
    with open(os.path.join(HERE, name + ".json")) as fh:
        meta = json.load(fh)
    return meta, Reader.open(os.path.join(HERE, name + ".ebiv")).packets()


@pytest.mark.parametrize("name", NAMES)
def test_plain_decode_matches_fixture(name):
    from govid_amd.codec import Codec

    meta, pkts = load(name)
    print(f"fixture {name}: {len(pkts)} packets", flush=True)
    c = Codec(ring=3)
    for i, p in enumerate(pkts):
        got = [zlib.crc32(a.tobytes()) for a in c.decode(p).planes()]
        assert got == meta["plane_crc32"][i], f"{name} frame {i}"
    c.close()


def test_batched_decode_matches_fixtures():
    """Every fixture at once, as sessions of one engine with a 4-frame window."""
    from govid_amd.codec import Batch

    b = Batch()
    want, sid_of = {}, {}
    for name in NAMES:
        meta, pkts = load(name)
        sid = b.open(ring=4, window=4)
        sid_of[sid] = name
        want[name] = meta["plane_crc32"]
        for i, p in enumerate(pkts):
            b.submit(sid, p, tag=i)
    n = 0
    while b.pending():
        for c in b.step():
            assert c.status == 0, c.error
            got = [zlib.crc32(a.tobytes()) for a in c.frame.planes()]
            assert got == want[sid_of[c.session]][c.tag], f"{sid_of[c.session]} frame {c.tag}"
            n += 1
    assert n == sum(len(v) for v in want.values())
    b.close()
