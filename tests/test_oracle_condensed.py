"""Pins the oracle's condensed-stream, per-cluster and 1/2-byte forms.

* condensed stream -- hand-derived streams (tests/golden/known_answers.json,
  "condensed"); the stream equals the concatenation of each row's slot cut at
  its size, the slot being pinned by the bit-list construction; offsets equal
  the running sum of the closed-form sizes; round trips against numpy
  indexing of the original;
* clusters         -- reduces row by row to the single-mask functions with the
  row's own mask row; all rows in one cluster c equals mask row c; no clusters
  equals mask row 0 (what the reference applies to every row);
* 1- and 2-byte    -- the bit-list brute force of the slot format at w = 1, 2.
"""
import json
import os

import numpy as np
import pytest

from test_oracle import bitlist_pack, bitlist_unpack

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")


def _hex(xs, w):
    return np.array([int(x, 16) for x in xs], dtype=np.dtype("u%d" % w))


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def _dataset(rng, n, v, w, zero_frac=0.4):
    """Rows with many elements that match a sparse mask: floats with zeros,
    viewed as w-byte words."""
    dt = np.dtype("u%d" % w)
    words = rng.integers(0, 1 << 62, size=(n, v), dtype=np.uint64).astype(dt)
    small = (rng.integers(0, 1 << 10, size=(n, v), dtype=np.uint64)).astype(dt)
    words = np.where(rng.random((n, v)) < 0.6, small, words)
    words[rng.random((n, v)) < zero_frac] = 0
    return words


def _kmeans_like_masks(oracle, x, k):
    """K mask rows from preprocess over K row groups (distinct, realistic)."""
    n = x.shape[0]
    masks, bitvals = [], []
    for c in range(k):
        part = x[c * n // k:(c + 1) * n // k]
        m, b, _, _ = oracle.preprocess(part, 0.75 + 0.05 * (c % 3))
        masks.append(m)
        bitvals.append(b)
    return np.stack(masks), np.stack(bitvals)


def test_condensed_hand_derived(oracle, golden):
    slots = {c["name"]: c for c in golden["slots"]}
    for case in golden["condensed"]:
        w = case["w"]
        rows, masks, bitvals = [], [], []
        for name in case["rows_from_slots"]:
            s = slots[name]
            rows.append(_hex(s["row"], w))
            masks.append(_hex(s["mask"], w))
            bitvals.append(_hex(s["bitval"], w))
        for r in case.get("extra_rows", []):
            rows.append(_hex(r, w))
        x = np.stack(rows)
        if case["clusters"] is None:
            mask, bitval, cl = masks[0], bitvals[0], None
        else:
            mask, bitval, cl = np.stack(masks), np.stack(bitvals), np.array(case["clusters"], dtype=np.int32)
        stream, offsets = oracle.compress_condensed(x, mask, bitval, clusters=cl)
        assert [int(v) for v in stream] == [int(v, 16) for v in case["stream"]], case["name"]
        assert offsets.tolist() == case["offsets"], case["name"]
        back = oracle.decompress_condensed(stream, offsets, x.shape[1], mask, bitval, clusters=cl)
        assert np.array_equal(back, x), case["name"]


@pytest.mark.parametrize("w", [1, 2, 4, 8])
@pytest.mark.parametrize("v", [1, 5, 9, 50, 64, 130])
def test_condensed_is_cut_slots(oracle, w, v):
    rng = np.random.default_rng(10 * w + v)
    x = _dataset(rng, 97, v, w)
    mask, bitval, _, _ = oracle.preprocess(x, 0.75)
    idx = rng.integers(0, 97, size=60)
    for index in (None, idx):
        stream, offsets = oracle.compress_condensed(x, mask, bitval, index=index)
        sizes, _ = oracle.get_compress_size(x, mask, bitval, index=index)
        assert np.array_equal(offsets, np.cumsum(sizes))
        slots, bm = oracle.compress(x, mask, bitval, index=index)
        raw = stream.view(np.uint8)
        start = 0
        for i in range(len(sizes)):
            row_bytes = slots[i].view(np.uint8)
            assert np.array_equal(raw[start:start + sizes[i]], row_bytes[:sizes[i]])
            # nothing is lost by the cut: the slot tail is zero
            assert not row_bytes[sizes[i]:].any()
            # a row is packed exactly when its stored length is below V*w
            assert bool((bm[i // 32] >> (i % 32)) & 1) == (sizes[i] < v * w)
            # independent construction of the same slot
            src = x[i if index is None else index[i]]
            words, size, packed = bitlist_pack(src, mask, bitval, w)
            assert size == sizes[i] and [int(t) for t in slots[i]] == words
            start += sizes[i]
        assert start * 1 == len(raw)
        want = x if index is None else x[index]
        assert np.array_equal(oracle.decompress_condensed(stream, offsets, v, mask, bitval), want)
        q = rng.integers(0, len(sizes), size=40)
        assert np.array_equal(oracle.decompress_condensed(stream, offsets, v, mask, bitval, index=q), want[q])


@pytest.mark.parametrize("w", [4, 8])
def test_clusters_reduce_to_single_mask(oracle, w):
    rng = np.random.default_rng(7 + w)
    n, v, k = 120, 40, 5
    x = _dataset(rng, n, v, w)
    masks, bitvals = _kmeans_like_masks(oracle, x, k)
    assert len({m.tobytes() for m in masks}) > 1  # the clusters really differ
    cl = rng.integers(0, k, size=n).astype(np.int32)
    sizes, ctr = oracle.get_compress_size(x, masks, bitvals, ctr=5, clusters=cl)
    slots, bm = oracle.compress(x, masks, bitvals, clusters=cl)
    for r in range(n):
        c = cl[r]
        s1, _ = oracle.get_compress_size(x[r:r + 1], masks[c], bitvals[c])
        sl1, bm1 = oracle.compress(x[r:r + 1], masks[c], bitvals[c])
        assert sizes[r] == s1[0]
        assert np.array_equal(slots[r], sl1[0])
        assert bool((bm[r // 32] >> (r % 32)) & 1) == bool(bm1[0] & 1)
    assert ctr == 5 + int(sizes.sum())
    assert np.array_equal(oracle.decompress_fetch(slots, masks, bitvals, bm, clusters=cl), x)
    q = rng.integers(0, n, size=77)
    assert np.array_equal(oracle.decompress_fetch(slots, masks, bitvals, bm, index=q, clusters=cl), x[q])
    # one cluster for every row == that mask row; no clusters == mask row 0
    for c in (0, k - 1):
        same = np.full(n, c, dtype=np.int32)
        a, abm = oracle.compress(x, masks, bitvals, clusters=same)
        b, bbm = oracle.compress(x, masks[c], bitvals[c])
        assert np.array_equal(a, b) and np.array_equal(abm, bbm)
    a, abm = oracle.compress(x, masks, bitvals)
    b, bbm = oracle.compress(x, masks[0], bitvals[0])
    assert np.array_equal(a, b) and np.array_equal(abm, bbm)
    # a planted wrong cluster id changes the result (the ids are really read)
    wrong = cl.copy()
    flip = rng.random(n) < 0.5
    wrong[flip] = (wrong[flip] + 1) % k
    out = oracle.decompress_fetch(slots, masks, bitvals, bm, clusters=wrong)
    assert not np.array_equal(out, x)


def test_clusters_with_index_are_source_rows(oracle):
    """Size and compress read clusters[index[i]] (the source row's cluster);
    the condensed fetch reads the cluster of the stored row."""
    rng = np.random.default_rng(3)
    n, v, k = 90, 33, 4
    x = _dataset(rng, n, v, 8)
    masks, bitvals = _kmeans_like_masks(oracle, x, k)
    cl = rng.integers(0, k, size=n).astype(np.int32)
    idx = rng.integers(0, n, size=50)
    sizes, _ = oracle.get_compress_size(x, masks, bitvals, index=idx, clusters=cl)
    direct, _ = oracle.get_compress_size(x[idx], masks, bitvals, clusters=cl[idx])
    assert np.array_equal(sizes, direct)
    stream, offsets = oracle.compress_condensed(x, masks, bitvals, index=idx, clusters=cl)
    s2, o2 = oracle.compress_condensed(x[idx], masks, bitvals, clusters=cl[idx])
    assert np.array_equal(stream, s2) and np.array_equal(offsets, o2)
    back = oracle.decompress_condensed(stream, offsets, v, masks, bitvals, clusters=cl[idx])
    assert np.array_equal(back, x[idx])


def _mask_small(w, k):
    B = 8 * w
    pats = [0, (1 << B) - 1, ((1 << B) - 1) ^ ((1 << (B // 2)) - 1), (1 << (B // 2)) - 1, 1 << (B - 1), 1,
            (1 << (B - 1)) | 1, int("10" * (B // 2), 2), ((1 << B) - 1) ^ 1, ((1 << B) - 1) >> 1]
    return pats[k % len(pats)]


@pytest.mark.parametrize("w", [1, 2])
@pytest.mark.parametrize("V", [1, 2, 3, 7, 8, 9, 16, 17, 33, 64, 65, 200])
def test_bitlist_brute_force_small_words(oracle, w, V):
    """The slot format at 8- and 16-bit words: mask classes x {flagged, raw}."""
    rng = np.random.default_rng(100 * w + V)
    dt = np.dtype("u%d" % w)
    B = 8 * w
    for trial in range(10):
        mask = np.array([_mask_small(w, trial + 3 * j) for j in range(V)], dtype=dt)
        bitval = rng.integers(0, 1 << B, size=V, dtype=np.uint64).astype(dt) & mask
        row = rng.integers(0, 1 << B, size=V, dtype=np.uint64).astype(dt)
        flag = rng.random(V) < 0.85
        row = np.where(flag, (row & ~mask) | bitval, row).astype(dt)
        words, size, packed = bitlist_pack(row, mask, bitval, w)
        comp, bm = oracle.compress(row[None, :], mask, bitval)
        assert oracle.row_size(row, mask, bitval) == size
        assert bool(bm[0] & 1) == packed
        assert [int(v) for v in comp[0]] == words
        assert bitlist_unpack(comp[0], mask, bitval, w, packed) == [int(v) for v in row]
        assert np.array_equal(oracle.decompress_fetch(comp, mask, bitval, bm)[0], row)


def test_compress_mt_matches_compress(oracle):
    """The multi-threaded host compress (timed by the bench) gives the same
    slots and bitmask as the single-threaded one, with an index and clusters."""
    rng = np.random.default_rng(12)
    n, v, k = 1000, 50, 3
    x = _dataset(rng, n, v, 8)
    masks, bitvals = _kmeans_like_masks(oracle, x, k)
    cl = rng.integers(0, k, size=n).astype(np.int32)
    idx = rng.integers(0, n, size=333)
    for kw in (dict(), dict(index=idx), dict(clusters=cl), dict(index=idx, clusters=cl)):
        m, b = (masks, bitvals) if "clusters" in kw else (masks[0], bitvals[0])
        a, abm = oracle.compress(x, m, b, **kw)
        for t in (1, 3, 8):
            c, cbm = oracle.compress_mt(x, m, b, t, **kw)
            assert np.array_equal(a, c) and np.array_equal(abm, cbm), (kw.keys(), t)
