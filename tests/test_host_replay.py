"""The decoder's host half (govid_amd/csrc/ebiv_host.cpp: frame header,
configure's checks, override split, tables and tile framing, and Go's state
transitions) built with g++ under AddressSanitizer and UndefinedBehaviorSanitizer
and replayed over the hostile inputs of the GPU error tests and a seeded
mutation corpus in the shape of fuzz_test.go:47-105.  Where the device would
decode tiles, the oracle's tile outcome is fed in; every packet's status must
equal the oracle's, and successful frames must take ring slots consistently
with the oracle's ring images (codec.go:84-88)."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import clip_kw, drop_front, edit_tables, flip_byte, zero_tile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ absent")
    exe = str(tmp_path_factory.mktemp("replay") / "host_replay")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "govid_amd", "csrc"),
           os.path.join(ROOT, "govid_amd", "csrc", "ebiv_host.cpp"), os.path.join(ROOT, "tests", "native", "host_replay.cpp"),
           "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def _run(exe, tmp_path, sequences, ring=3):
    """sequences: lists of packets (bytes) or "flush"; each on a new codec.
    Returns (harness lines, oracle outcomes) per sequence."""
    blob = bytearray(struct.pack("<I", ring))
    want = []
    for seq in sequences:
        blob.append(2)
        dec = O.OracleCodec(ring=ring)
        out = []
        for p in seq:
            if p == "flush":
                blob.append(1)
                dec.flush()
                continue
            st, fr = dec.decode_raw(p)
            stage = dec.last_tile_failure() if st else 0
            out.append((st, fr.y if st == 0 else None))
            blob += bytes([0]) + struct.pack("<I", len(p)) + p + bytes([stage != 1, stage != 2])
        dec.close()
        want.append(out)
    path = tmp_path / "seq.bin"
    path.write_bytes(bytes(blob))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(path)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    lines = [tuple(map(int, ln.split())) for ln in r.stdout.splitlines()]
    got, k = [], 0
    for out in want:
        got.append(lines[k:k + len(out)])
        k += len(out)
    assert k == len(lines)
    return got, want


def _check(got, want, labels):
    for g, w, label in zip(got, want, labels):
        slot_of = {}
        for i, ((gs, gslot), (ws, wy)) in enumerate(zip(g, w)):
            assert gs == ws, f"{label} packet {i}: host half says {gs}, oracle {ws}"
            if ws == 0:  # the same ring image for the same slot
                assert slot_of.setdefault(gslot, wy) == wy, f"{label} packet {i}: ring slot {gslot}"


def _mutate(rng, p: bytes) -> bytes:
    b = bytearray(p)
    kind = int(rng.integers(6))
    if kind == 0 and b:
        for _ in range(int(rng.integers(1, 9))):
            i = int(rng.integers(len(b)))
            b[i] ^= 1 << int(rng.integers(8))
    elif kind == 1 and b:
        b = b[:int(rng.integers(len(b)))]
    elif kind == 2:
        i = int(rng.integers(len(b) + 1))
        b[i:i] = rng.integers(0, 256, int(rng.integers(1, 9))).astype(np.uint8).tobytes()
    elif kind == 3 and len(b) > 1:
        i = int(rng.integers(len(b)))
        del b[i:i + int(rng.integers(1, 17))]
    elif kind == 4 and len(b) > 8:  # bytes in the header and framing region
        for _ in range(int(rng.integers(1, 4))):
            b[int(rng.integers(min(len(b), 40)))] = int(rng.integers(256))
    else:
        i = int(rng.integers(len(b) + 1))
        n = int(rng.integers(1, 33))
        b[i:i + n] = rng.integers(0, 256, n).astype(np.uint8).tobytes()
    return bytes(b)


def test_hostile_packets_match_the_oracle(replay, tmp_path):
    pk = list(clip_kw("synthpan", 64, 48, 6, qp=22, gop=3).packets)
    other = clip_kw("synthpan", 80, 48, 1, qp=22, gop=1).packets[0]
    ov = list(clip_kw("parallax", 256, 192, 10, qp=20, gop=5, tile_cols=2, tile_rows=2, altref=4, threads=4).packets)
    big = bytes([0x04, 0x01, 0x20, 0xE0, 0x10, 0]) + bytes(64)  # 8193x4320
    seqs = {
        "bad packets": [b"", bytes([1]), bytes([0, 1, 0]), big, bytes([0x1D]), pk[0], bytes([0x1D]) + pk[1][1:],
                        bytes([0x21]) + pk[1][1:], pk[1], other, pk[2], "flush", pk[3], pk[4]],
        "tile failures": [pk[0], zero_tile(pk[1]), pk[2], pk[3], drop_front(pk[4], 0, 0.5), pk[5]],
        "override failures": [ov[0], drop_front(ov[1], 0, 0.5), ov[2], drop_front(ov[5], 1, 0.5, 1), ov[6],
                              drop_front(ov[5], 2, 0.5, 0), ov[7], flip_byte(ov[0], 0, 0.3, 1), ov[1]],
        "tables": [edit_tables(pk[0], lambda e: [(1, []) if m == 2 else (m, v) for m, v in e]), pk[0],
                   edit_tables(pk[1], lambda e: [(2, [4095] + [0] * (len(v) - 1)) if m == 2 else (m, v)
                                                 for m, v in e]), pk[1]],
    }
    got, want = _run(replay, tmp_path, list(seqs.values()))
    _check(got, want, list(seqs))


def test_seeded_mutations_match_the_oracle(replay, tmp_path):
    """600 mutated packets, cold and after a key frame, and 40 streams with
    several mutated packets each, replayed in one sanitized process."""
    seeds = [list(clip_kw("synthframe", 64, 48, 6, qp=20, gop=4, altref=4).packets),
             list(clip_kw("rects", 96, 64, 5, qp=22, gop=3, tile_cols=2, tile_rows=2).packets),
             list(clip_kw("synthpan", 37, 21, 4, qp=22, gop=2, coding="raw").packets)]
    rng = np.random.default_rng(105)
    seqs, labels = [], []
    for case in range(300):
        pk = seeds[case % len(seeds)]
        m = _mutate(rng, pk[int(rng.integers(len(pk)))])
        seqs += [[m], [pk[0], m, pk[1 % len(pk)]]]
        labels += [f"cold {case}", f"warm {case}"]
    for case in range(40):
        pk = list(seeds[case % len(seeds)])
        for _ in range(3):
            i = int(rng.integers(len(pk)))
            pk[i] = _mutate(rng, pk[i])
        if case % 4 == 0:
            pk.insert(int(rng.integers(1, len(pk))), "flush")
        seqs.append(pk)
        labels.append(f"stream {case}")
    got, want = _run(replay, tmp_path, seqs)
    _check(got, want, labels)
    statuses = {s for w in want for s, _ in w}
    assert {0, 1}.issubset(statuses), statuses
