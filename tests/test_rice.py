"""The compressed pair stream (the reference's record layout, pbig.cu:849-888,
decode.cc:26-138): round trips of the host encoder/decoder and decoding of
streams built byte by byte in the reference's layout.  No GPU."""
import struct

import numpy as np
import pytest

UNUSED = 0x0FFFFF11


def canon(p):
    p = np.asarray(p, dtype=np.uint32).reshape(-1, 2)
    lo, hi = np.minimum(p[:, 0], p[:, 1]), np.maximum(p[:, 0], p[:, 1])
    k = (hi.astype(np.uint64) << np.uint64(32)) | lo.astype(np.uint64)
    return np.sort(k)


def record(seg1, seg2, l_rid, offsets):
    """One record as the reference's encoder lays it out, padded to words."""
    b = struct.pack("<BBHH", seg1, seg2, l_rid, len(offsets)) + bytes(offsets)
    b += b"\0" * (-len(b) % 8)
    return list(np.frombuffer(b, dtype="<u8"))


def test_decode_stream_in_reference_layout():
    import pbig_amd

    recs = np.arange(1000, 1000 + 700, dtype=np.uint32)  # cell 0: entries 0..599, cell 1: 600..699
    start = np.array([0, 600, 700], dtype=np.uint32)
    words = [UNUSED, 0] + record(1, 0, 5, [2, 7, 9]) + record(2, 1, 0, [255]) + \
            [UNUSED, 1] + record(0, 0, 99, list(range(10)))
    got = pbig_amd.rice_decode(np.array(words, dtype=np.uint64), start, recs)
    want = [(1000 + 256 + 5, 1000 + 2), (1000 + 256 + 5, 1000 + 7), (1000 + 256 + 5, 1000 + 9),
            (1000 + 512, 1000 + 256 + 255)] + [(1600 + 99, 1600 + t) for t in range(10)]
    assert np.array_equal(canon(got), canon(want))


@pytest.mark.parametrize("n,density,seed", [(50, 0.3, 1), (2000, 0.01, 2), (600, 0.9, 3), (1, 0.0, 4)])
def test_round_trip_random(n, density, seed):
    import pbig_amd

    rng = np.random.default_rng(seed)
    i, j = np.triu_indices(n, 1)
    keep = rng.random(len(i)) < density
    pairs = np.stack([i[keep], j[keep]], 1).astype(np.uint32)
    rng.shuffle(pairs)
    flip = rng.random(len(pairs)) < 0.5
    pairs[flip] = pairs[flip][:, ::-1]  # either orientation is accepted
    w, s, r = pbig_amd.rice_encode(pairs, n)
    assert len(s) >= 1 and s[-1] == len(r)
    assert np.array_equal(canon(pbig_amd.rice_decode(w, s, r)), canon(pairs))


def test_round_trip_many_partners_spans_segments_and_cells():
    """A box with 70,000 partners needs 256-entry segments and two cells."""
    import pbig_amd

    n = 70001
    pairs = np.stack([np.zeros(n - 1, np.uint32), np.arange(1, n, dtype=np.uint32)], 1)
    pairs = np.concatenate([pairs, [[5, 6], [7, 69999]]]).astype(np.uint32)
    w, s, r = pbig_amd.rice_encode(pairs, n)
    assert len(s) - 1 >= 4  # box 0 in two cells, boxes 5 and 7 in one each
    assert np.array_equal(canon(pbig_amd.rice_decode(w, s, r)), canon(pairs))
    # about one byte per pair in the stream
    assert len(w) * 8 < 1.2 * len(pairs) + 16 * (len(s) - 1) + 8 * 300


def test_empty_and_malformed():
    import pbig_amd

    w, s, r = pbig_amd.rice_encode(np.zeros((0, 2), np.uint32), 10)
    assert len(w) == 0 and len(pbig_amd.rice_decode(w, s, r)) == 0
    with pytest.raises(pbig_amd.PbigError):
        pbig_amd.rice_encode(np.array([[3, 3]], np.uint32), 10)  # self pair
    with pytest.raises(pbig_amd.PbigError):
        pbig_amd.rice_encode(np.array([[3, 12]], np.uint32), 10)  # id past n_boxes
    start = np.array([0, 2], np.uint32)
    recs = np.array([4, 5], np.uint32)
    with pytest.raises(pbig_amd.PbigError):  # no block header
        pbig_amd.rice_decode(np.array(record(0, 0, 0, [1]), np.uint64), start, recs)
    with pytest.raises(pbig_amd.PbigError):  # partner index past the cell lists
        pbig_amd.rice_decode(np.array([UNUSED, 0] + record(0, 0, 0, [9]), np.uint64), start, recs)
