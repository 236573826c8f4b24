"""The TRF1 transfer-cache framing (hl2sdk_tools_sharp_amd.trf1) against the
reference's codec tests (Tests/MapTools/Rad/Bounce/TransferSetCodecTests.cs):
a set survives index, chunks and packing entry for entry, the chunks are cut
as the reference cuts them, and what does not fit is refused."""
from __future__ import annotations

import struct

import numpy as np
import pytest

from hl2sdk_tools_sharp_amd import trf1

EPS = np.float32(np.finfo(np.float32).smallest_subnormal)  # float.Epsilon


def sample():
    """The reference's Sample(): four patches with 3, 0, 5 and 2 transfers,
    patch numbers going up and down (TransferSetCodecTests.cs:22-33), in the
    order For(p) yields them."""
    lists = [
        [(1, 0.5), (2, 0.25), (3, 1e-7)],
        [],
        [(0, 0.1), (3, 0.2), (1, 0.3), (2, EPS), (0, 0.9)],
        [(0, 0.75), (2, 0.125)],
    ]
    counts = np.array([len(x) for x in lists], np.int32)
    flat = [t for x in lists for t in x]
    rec = np.zeros((len(flat), 2), "<u4")
    rec[:, 0] = [p for p, _ in flat]
    rec[:, 1] = np.array([w for _, w in flat], np.float32).view("<u4")
    return counts, rec.tobytes(), lists


def as_lists(counts, src, w):
    out, at = [], 0
    for c in counts:
        out.append([(int(src[at + i]), np.float32(w[at + i])) for i in range(c)])
        at += c
    return out


def same(a, b):
    return len(a) == len(b) and all(
        len(x) == len(y) and all(p == q and np.float32(u).view(np.uint32) == np.float32(v).view(np.uint32)
                                 for (p, u), (q, v) in zip(x, y)) for x, y in zip(a, b))


def test_index_layout():
    counts, _, _ = sample()
    idx = trf1.index(counts)
    assert idx[:4] == b"TRF1"
    assert struct.unpack_from("<iiq", idx, 4) == (4, 5, 10)
    assert list(np.frombuffer(idx, "<i4", offset=20)) == [3, 0, 5, 2]


def test_a_set_round_trips_through_index_and_chunks():
    counts, rec, lists = sample()
    back = trf1.read(trf1.index(counts), trf1.chunks(rec, 1 << 20), 4)
    assert back is not None
    c, src, w, mx = back
    assert same(as_lists(c, src, w), lists)
    assert mx == 5 and int(c.sum()) == 10


def test_chunks_are_whole_transfers_and_cover_the_set():
    counts, rec, lists = sample()
    ch = trf1.chunks(rec, 24)
    assert [len(x) for x in ch] == [24, 24, 24, 8]
    c, src, w, _ = trf1.read(trf1.index(counts), ch, 4)
    assert same(as_lists(c, src, w), lists)
    assert [len(x) for x in trf1.chunks(rec, 3)] == [8] * 10  # at least one transfer a chunk


def test_pack_round_trips_every_chunk():
    _, rec, _ = sample()
    for chunk in trf1.chunks(rec, 24):
        assert trf1.unpack(trf1.pack(chunk)) == chunk
    # a large chunk with patch numbers jumping both ways
    rng = np.random.default_rng(1)
    big = np.zeros((5000, 2), "<u4")
    big[:, 0] = rng.integers(0, 2**31 - 1, 5000)
    big[:, 1] = rng.standard_normal(5000).astype(np.float32).view("<u4")
    assert trf1.unpack(trf1.pack(big.tobytes())) == big.tobytes()


def test_an_empty_set_round_trips():
    counts = np.zeros(3, np.int32)
    back = trf1.read(trf1.index(counts), trf1.chunks(b"", 24), 3)
    assert back is not None and int(back[0].sum()) == 0 and back[3] == 0


def test_what_does_not_fit_is_refused():
    counts, rec, _ = sample()
    idx = trf1.index(counts)
    ch = trf1.chunks(rec, 24)
    assert trf1.read(idx, trf1.chunks(rec, 1 << 20), 5) is None  # another patch count
    assert trf1.read(idx, ch[:-1], 4) is None  # a chunk missing
    bad = bytearray(rec)
    struct.pack_into("<i", bad, 0, 7)  # a patch out of range
    assert trf1.read(idx, [bytes(bad)], 4) is None
    assert trf1.read(idx, [rec[:-1]], 4) is None  # not whole transfers
    wrong_max = bytearray(idx)
    struct.pack_into("<i", wrong_max, 8, 4)
    assert trf1.read(bytes(wrong_max), [rec], 4) is None
    neg = bytearray(idx)
    struct.pack_into("<i", neg, 24, -1)
    assert trf1.read(bytes(neg), [rec], 4) is None
    assert trf1.read(b"TRF0" + idx[4:], [rec], 4) is None


@pytest.mark.parametrize("junk", [bytes([1, 2]), bytes([0xFF] * 4), bytes([10, 0, 0, 0, 0xDE, 0xAD, 0xBE, 0xEF])])
def test_junk_does_not_unpack(junk):
    assert trf1.unpack(junk) is None


def test_truncated_or_trailing_streams_do_not_unpack():
    _, rec, _ = sample()
    p = trf1.pack(rec)
    assert trf1.unpack(p[:-3]) is None
    assert trf1.unpack(p + b"\x00") is None
    wrong = bytearray(p)
    struct.pack_into("<i", wrong, 0, 9)  # a count the stream does not hold
    assert trf1.unpack(bytes(wrong)) is None
