"""Pins the CPU oracle (oracle/ibp_oracle.c) before anything is checked against it.

The reference publishes no golden vectors (its only check is a printed round
trip, tests/compress_test.py:21), so each oracle function is pinned here by
something other than its own formula:

* slot format   -- hand-derived slots (tests/golden/known_answers.json) and an
                   independent bit-list construction of the format;
* size          -- the closed form against the length of the packed stream;
* unpack/gather -- round trips and numpy fancy indexing of the original;
* preprocess    -- hand-derived cases, a numpy brute force (np.unpackbits
                   counting, float32/float64 comparisons), special cases;
* kmeans        -- glibc rand() known answers, libstdc++ std::random_shuffle
                   compiled with g++, a vectorised numpy restatement and
                   separable-data known answers.
"""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")


def _hex_arr(xs, w):
    return np.array([int(x, 16) for x in xs], dtype=np.dtype("u%d" % w))


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


# ---------------------------------------------------------------- slot format

def bitlist_pack(row, mask, bitval, w):
    """Independent construction of a slot as explicit lists of bits."""
    B = 8 * w
    V = len(row)
    H = -(-(-(-V // 8)) // w) * w  # align_w(ceil(V/8)) in bytes
    header = [0] * (H * 8)
    stream = []
    S = 0
    for j in range(V):
        x, m, b = int(row[j]), int(mask[j]), int(bitval[j])
        if (x & m) == b:
            header[(j // B) * B + (j % B)] = 1
            S += bin(m).count("1")
            stream += [(x >> k) & 1 for k in range(B - 1, -1, -1) if not (m >> k) & 1]
        else:
            stream += [(x >> k) & 1 for k in range(B - 1, -1, -1)]
    size = min(V * w, H + -(-(V * w - S // 8) // w) * w)
    if size >= V * w:
        return [int(v) for v in row], size, False
    words = [0] * V
    for i, bit in enumerate(header):
        words[i // B] |= bit << (i % B)
    for p, bit in enumerate(stream):
        words[H // w + p // B] |= bit << (B - 1 - p % B)
    return words, size, True


def bitlist_unpack(slot, mask, bitval, w, packed):
    B = 8 * w
    V = len(slot)
    if not packed:
        return [int(v) for v in slot]
    H = -(-(-(-V // 8)) // w) * w
    bits = []
    for word in slot[H // w:]:
        bits += [(int(word) >> k) & 1 for k in range(B - 1, -1, -1)]
    out, p = [], 0
    for j in range(V):
        m = int(mask[j])
        if (int(slot[j // B]) >> (j % B)) & 1:
            v = int(bitval[j])
            for k in range(B - 1, -1, -1):
                if not (m >> k) & 1:
                    v |= bits[p] << k
                    p += 1
        else:
            v = 0
            for k in range(B - 1, -1, -1):
                v |= bits[p] << k
                p += 1
        out.append(v)
    return out


def test_hand_derived_slots(oracle, golden):
    for case in golden["slots"]:
        w = case["w"]
        row, mask, bitval = (_hex_arr(case[k], w) for k in ("row", "mask", "bitval"))
        want = _hex_arr(case["slot"], w)
        assert oracle.row_size(row, mask, bitval) == case["size"], case["name"]
        comp, bm = oracle.compress(row[None, :], mask, bitval)
        assert bool(bm[0] & 1) == case["packed"], case["name"]
        assert np.array_equal(comp[0], want), (case["name"], [hex(v) for v in comp[0]])
        if case["packed"]:
            assert np.array_equal(oracle.pack_row(row, mask, bitval), want), case["name"]
            assert np.array_equal(oracle.unpack_row(want, mask, bitval), row), case["name"]
        # the independent bit-list construction agrees with the hand derivation
        words, size, packed = bitlist_pack(row, mask, bitval, w)
        assert size == case["size"] and packed == case["packed"]
        assert words == [int(v) for v in want], case["name"]


MASK_CLASSES_64 = [0, (1 << 64) - 1, 0xFFFFFFFF00000000, 0x00000000FFFFFFFF, 0xF0F0F0F0F0F0F0F0,
                   0xAAAAAAAAAAAAAAAA, 1 << 63, 1, 0x8000000000000001, 0x7FFFFFFFFFFFFFFE,
                   0xFFF0000000000000, 0x000FFFFFFFFFFFFF]


def _mask_class(w, k):
    B = 8 * w
    m = MASK_CLASSES_64[k % len(MASK_CLASSES_64)]
    if B == 32:
        m = (m >> 32) | (m & 0xFFFFFFFF) if k % 3 == 0 else m & 0xFFFFFFFF
    return m & ((1 << B) - 1)


@pytest.mark.parametrize("w", [4, 8])
@pytest.mark.parametrize("V", [1, 2, 3, 4, 7, 8, 9, 33, 64, 65])
def test_bitlist_brute_force(oracle, w, V):
    """Mask classes {0, ~0, single runs, alternating, MSB/LSB} x {flagged, raw}."""
    rng = np.random.default_rng(1000 * w + V)
    dt = np.dtype("u%d" % w)
    B = 8 * w
    for trial in range(12):
        mask = np.array([_mask_class(w, trial + 5 * j) for j in range(V)], dtype=dt)
        rnd = rng.integers(0, 1 << 62, size=V, dtype=np.uint64).astype(dt) ^ rng.integers(
            0, 4, size=V, dtype=np.uint64).astype(dt) << dt.type(B - 2)
        bitval = rnd & mask
        row = rng.integers(0, 1 << 62, size=V, dtype=np.uint64).astype(dt)
        flag = rng.random(V) < 0.8
        row = np.where(flag, (row & ~mask) | bitval, row)
        words, size, packed = bitlist_pack(row, mask, bitval, w)
        comp, bm = oracle.compress(row[None, :], mask, bitval)
        assert oracle.row_size(row, mask, bitval) == size
        assert bool(bm[0] & 1) == packed
        assert [int(v) for v in comp[0]] == words
        assert bitlist_unpack(comp[0], mask, bitval, w, packed) == [int(v) for v in row]
        out = oracle.decompress_fetch(comp, mask, bitval, bm)
        assert np.array_equal(out[0], row)


@pytest.mark.parametrize("w", [4, 8])
def test_size_closed_form_vs_stream_length(oracle, w):
    """s(r) from the closed form equals header + aligned payload length of the
    stream the packer produces, and obeys s <= V*w, s % w == 0."""
    rng = np.random.default_rng(7 + w)
    dt = np.dtype("u%d" % w)
    B = 8 * w
    for V in (5, 64, 100):
        data = rng.integers(0, 1 << 62, size=(50, V), dtype=np.uint64).astype(dt)
        data &= dt.type((1 << (B - 6)) - 1)
        data[::3, ::2] = 0
        mask, bitval, _, _ = oracle.preprocess(data, 0.8)
        sizes, _ = oracle.get_compress_size(data, mask, bitval)
        H = -(-(-(-V // 8)) // w) * w
        for r in range(data.shape[0]):
            flags = (data[r] & mask) == bitval
            bits = int(sum(B - bin(int(m)).count("1") if f else B for f, m in zip(flags, mask)))
            stream_bytes = H + -(-(-(-bits // 8)) // w) * w
            assert sizes[r] == min(V * w, stream_bytes)
            assert sizes[r] <= V * w and sizes[r] % w == 0


def test_size_index_and_counter(oracle):
    rng = np.random.default_rng(3)
    data = (rng.integers(0, 1 << 20, size=(40, 16), dtype=np.uint64))
    mask = np.full(16, 0xFFFFFFFFFFF00000, dtype=np.uint64)
    bitval = np.zeros(16, dtype=np.uint64)
    full, _ = oracle.get_compress_size(data, mask, bitval)
    idx = np.array([5, 0, 39, 5, 12])
    sub, ctr = oracle.get_compress_size(data, mask, bitval, idx, ctr=100)
    assert np.array_equal(sub, full[idx])
    assert ctr == 100 + int(full[idx].sum())  # added to, never reset


@pytest.mark.parametrize("w", [4, 8])
def test_gather_matches_fancy_indexing(oracle, w):
    rng = np.random.default_rng(11)
    dt = np.dtype("u%d" % w)
    data = rng.integers(0, 1 << 62, size=(300, 37), dtype=np.uint64).astype(dt)
    data &= dt.type((1 << (8 * w - 9)) - 1)
    data[rng.random(data.shape) < 0.3] = 0
    mask, bitval, _, _ = oracle.preprocess(data)
    comp, bm = oracle.compress(data, mask, bitval)
    idx = rng.integers(0, 300, size=500)  # duplicates allowed
    out = oracle.decompress_fetch(comp, mask, bitval, bm, idx)
    assert np.array_equal(out, data[idx])
    # in-place compress gives the same slots and bitmask as out-of-place
    inplace = data.copy()
    bm2 = oracle.compress_inplace(inplace, mask, bitval)
    assert np.array_equal(inplace, comp) and np.array_equal(bm, bm2)


def test_compress_with_index(oracle):
    rng = np.random.default_rng(5)
    data = rng.integers(0, 1 << 10, size=(64, 20), dtype=np.uint64)
    mask = np.full(20, ~np.uint64(0) << np.uint64(12), dtype=np.uint64)
    bitval = np.zeros(20, dtype=np.uint64)
    idx = np.array([3, 3, 60, 1, 0])
    comp, bm = oracle.compress(data, mask, bitval, idx)
    for i, r in enumerate(idx):
        single, sbm = oracle.compress(data[r:r + 1], mask, bitval)
        assert np.array_equal(comp[i], single[0]) and bool(bm[i // 32] >> (i % 32) & 1) == bool(sbm[0] & 1)


# ---------------------------------------------------------------- preprocess

def numpy_preprocess(data, threshold):
    """Brute force of A.1 with numpy: unpackbits counting, float32 / float64
    comparisons, the float32 sweep, first strict maximum, zero default."""
    n, v = data.shape
    w = data.itemsize
    B = 8 * w
    bits = np.unpackbits(data.view(np.uint8).reshape(n, v, w)[..., ::-1], axis=2)  # MSB..LSB per element
    cnt = bits.sum(axis=0)[:, ::-1].astype(np.int64)  # [v, B], index = bit position
    if threshold is None:
        ts, t = [], np.float32(0.7)
        while t <= np.float32(1.0):
            ts.append(t)
            t = np.float32(np.float64(t) + 0.05)
    else:
        ts = [np.float32(threshold)]
    best, out_m, out_b, step = 0.0, np.zeros(v, np.uint64), np.zeros(v, np.uint64), -1
    pw = (np.uint64(1) << np.arange(B, dtype=np.uint64))
    for s, t in enumerate(ts):
        hi = np.float32(t) * np.float32(n)
        lo = (1.0 - np.float64(t)) * np.float64(n)
        gt = cnt.astype(np.float32) > hi
        lt = (~gt) & (cnt.astype(np.float64) < lo)
        m = ((gt | lt) * pw).sum(axis=1, dtype=np.uint64)
        b = (gt * pw).sum(axis=1, dtype=np.uint64)
        du = data.astype(np.uint64) if w < 8 else data.view(np.uint64)
        flag = (du & m) == b
        pc = np.array([bin(int(x)).count("1") for x in m])
        ctr = (flag * pc).sum(axis=1)
        saved = int(np.where(ctr > v, ctr - v, 0).sum())
        if float(saved) > best:
            best, out_m, out_b, step = float(saved), m, b, s
    dt = np.dtype("u%d" % w)
    return out_m.astype(dt), out_b.astype(dt), step


def test_preprocess_hand_derived(oracle, golden):
    for case in golden["preprocess"]:
        w = case["w"]
        rows = np.array([[int(x, 16) for x in r] for r in case["rows"]], dtype=np.dtype("u%d" % w))
        m, b, step, _ = oracle.preprocess(rows, case["threshold"])
        assert [hex(x) for x in m] == [hex(int(x, 16)) for x in case["mask"]], case["name"]
        assert [hex(x) for x in b] == [hex(int(x, 16)) for x in case["bitval"]], case["name"]
        assert step == case["step"], case["name"]


@pytest.mark.parametrize("w", [1, 2, 4, 8])
@pytest.mark.parametrize("threshold", [None, 0.9, 0.75])
def test_preprocess_numpy_brute_force(oracle, w, threshold):
    rng = np.random.default_rng(w * 10 + (0 if threshold is None else int(threshold * 100)))
    dt = np.dtype("u%d" % w)
    n, v = 257, 9
    f = rng.normal(0, 0.1, size=(n, v * w)).astype(np.float32)
    f[rng.random(f.shape) < 0.3] = 0.0
    data = np.ascontiguousarray(f.view(np.uint8).reshape(n, -1)[:, : v * w]).view(dt)
    m, b, step, _ = oracle.preprocess(data, threshold)
    m2, b2, step2 = numpy_preprocess(data, threshold)
    assert np.array_equal(m, m2) and np.array_equal(b, b2) and step == step2


def test_preprocess_special_cases(oracle):
    # constant column: mask all ones, bitval = the value
    data = np.tile(np.array([0x0123456789ABCDEF, 7], dtype=np.uint64), (20, 1))
    data[:, 1] = np.arange(20)
    m, b, _, _ = oracle.preprocess(data, 0.9)
    assert m[0] == np.uint64(0xFFFFFFFFFFFFFFFF) and b[0] == np.uint64(0x0123456789ABCDEF)
    # sweep: 7 float32 steps, the last exactly 1.0
    t, steps = oracle.sweep_last()
    assert steps == 7 and t == 1.0


def test_count_bits_matches_unpackbits(oracle):
    rng = np.random.default_rng(2)
    for w in (1, 2, 4, 8):
        dt = np.dtype("u%d" % w)
        data = rng.integers(0, 255, size=(33, 5 * w), dtype=np.uint8).view(dt)
        cnt = oracle.count_bits(data).reshape(data.shape[1], 8 * w)
        bits = np.unpackbits(data.view(np.uint8).reshape(33, -1, w), axis=2, bitorder="little")
        assert np.array_equal(cnt, bits.sum(axis=0))


# ---------------------------------------------------------------- kmeans

def test_glibc_rand_known_answers(golden):
    import ctypes
    libc = ctypes.CDLL(None)
    libc.srand(1)
    assert [libc.rand() for _ in range(10)] == golden["glibc_rand_seed1"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_random_shuffle_matches_libstdcxx(oracle):
    src = r"""
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>
int main(int argc, char** argv) {
    int n = atoi(argv[1]); srand(atoi(argv[2]));
    std::vector<int> v(n); for (int i = 0; i < n; ++i) v[i] = i;
    std::random_shuffle(v.begin(), v.end());
    for (int i = 0; i < n; ++i) printf("%d ", v[i]);
    return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "s.cpp"), "w") as f:
            f.write(src)
        exe = os.path.join(d, "s")
        subprocess.run(["g++", "-std=c++11", "-w", "-O1", "-o", exe, os.path.join(d, "s.cpp")], check=True)
        for n, seed in ((1, 1), (10, 1), (1000, 7), (4096, 1)):
            got = subprocess.run([exe, str(n), str(seed)], check=True, capture_output=True, text=True).stdout
            want = [int(x) for x in got.split()]
            assert list(oracle.random_shuffle(n, seed=seed)) == want


def numpy_kmeans(data, K, init_rows):
    """Vectorised restatement of A.7 (distances by popcount of XOR, majority
    bits with the stride-32 OR accumulation, final threshold-1.0 masks)."""
    n, v = data.shape
    w = data.itemsize
    B = 8 * w
    bits = np.unpackbits(data.view(np.uint8).reshape(n, v, w), axis=2, bitorder="little").astype(np.int64)
    cent_bits = bits[init_rows].copy()  # [K, v, B]
    if K > len(init_rows):
        cent_bits = np.concatenate([cent_bits, np.zeros((K - len(init_rows), v, B), np.int64)])
    for _ in range(10):
        dist = (bits[:, None, :, :] != cent_bits[None, :, :, :]).sum(axis=(2, 3))  # [n, K]
        cl = K - 1 - np.argmin(dist[:, ::-1], axis=1)  # highest index among minima
        cnt = np.bincount(cl, minlength=K)
        bs = np.zeros((K, v, B), np.int64)
        np.add.at(bs, cl, bits)
        maj = bs > (cnt // 2)[:, None, None]
        acc = maj.copy()
        for j in range(32, v):
            acc[:, j] |= acc[:, j - 32]
        cent_bits = acc.astype(np.int64)
    dist = None
    cnt = np.bincount(cl, minlength=K)
    bs = np.zeros((K, v, B), np.int64)
    np.add.at(bs, cl, bits)
    hi = bs.astype(np.float32) >= (np.float32(1.0) * cnt.astype(np.float32))[:, None, None]
    lo = (~hi) & (bs.astype(np.float64) <= 0.0 * cnt[:, None, None])
    pw = (np.uint64(1) << np.arange(B, dtype=np.uint64))
    dt = np.dtype("u%d" % w)
    masks = ((hi | lo) * pw).sum(axis=2, dtype=np.uint64).astype(dt)
    vals = (hi * pw).sum(axis=2, dtype=np.uint64).astype(dt)
    return masks, vals, cl.astype(np.int32)


@pytest.mark.parametrize("w,V,K", [(4, 40, 4), (8, 70, 3), (4, 5, 6)])
def test_kmeans_vs_numpy(oracle, w, V, K):
    rng = np.random.default_rng(w + V + K)
    dt = np.dtype("u%d" % w)
    n = 60
    base = rng.integers(0, 1 << 62, size=(4, V), dtype=np.uint64).astype(dt)
    data = base[rng.integers(0, 4, size=n)] ^ (rng.integers(0, 16, size=(n, V), dtype=np.uint64).astype(dt))
    seed = 3
    perm = oracle.random_shuffle(n, seed=seed)
    init = [0] + [int(perm[i]) for i in range(1, min(K, n))]
    m, b, cl = oracle.kmeans(data, K, seed=seed)
    m2, b2, cl2 = numpy_kmeans(data, K, init)
    assert np.array_equal(cl, cl2)
    assert np.array_equal(m, m2) and np.array_equal(b, b2)


def test_kmeans_separable_known_answer(oracle):
    """Two groups (zeros / ones in the high bits, 3 noisy low bits): with an
    initial centroid in each group, clusters are the groups and each mask is
    every bit except the noisy ones, bitval the group's constant bits."""
    n, v = 40, 8
    rng = np.random.default_rng(0)
    group = np.zeros(n, dtype=np.int64)
    group[1::2] = 1
    data = np.where(group[:, None] == 1, np.uint64(0xFFFFFFFFFFFFFFF8), np.uint64(0))
    data = data | rng.integers(0, 8, size=(n, v), dtype=np.uint64)
    seed = next(s for s in range(1, 100) if group[oracle.random_shuffle(n, seed=s)[1]] == 1)
    m, b, cl = oracle.kmeans(data, 2, seed=seed)
    assert np.array_equal(cl, group.astype(np.int32))
    assert np.all(m == np.uint64(0xFFFFFFFFFFFFFFF8))
    assert np.all(b[0] == 0) and np.all(b[1] == np.uint64(0xFFFFFFFFFFFFFFF8))


def test_kmeans_empty_cluster_all_ones(oracle):
    """An empty cluster gets mask = bitval = all ones (0 >= 1.0 * 0 holds)."""
    n, v = 12, 4
    data = np.zeros((n, v), dtype=np.uint32)
    data[6:] = 0xFFFFFFFF
    m, b, cl = oracle.kmeans(data, 5, seed=1)
    used = set(cl.tolist())
    for c in range(5):
        if c not in used:
            assert np.all(m[c] == 0xFFFFFFFF) and np.all(b[c] == 0xFFFFFFFF)
    assert len(used) <= 2


# Rows of the stride-32 OR-chain known answer: 3 rows of 33 int32 words, zero
# except columns 0 and 1 (A = 0x0000FFFF, 16 bits; D = 0xFF000000, 8 bits).
OR_CHAIN_ROWS = ((0x0000FFFF, 0), (0x0000FFFF, 0xFF000000), (0, 0xFF000000))


def or_chain_data():
    x = np.zeros((3, 33), dtype=np.uint32)
    for r, (c0, c1) in enumerate(OR_CHAIN_ROWS):
        x[r, 0], x[r, 1] = c0, c1
    return x.view(np.int32)


def or_chain_expected():
    """Hand derivation (ibp_preproc_kmeans.cuh:220-236; host :150-233), K = 2.
    srand(1): random_shuffle(0..2) = [0, 2, 1], so the initial centroids are
    row 0 = (A, 0, .., 0) and row 2 = (0, D, 0, ..).  Hamming distances
    (d0, d1): row0 (0, 24) row1 (8, 16) row2 (24, 0) -> clusters [0, 0, 1].
    construct_bitmasks: lane 0 visits columns 0 and 32 with one accumulator,
    so centroid 0 becomes (A, 0, .., 0, A) -- column 32 inherits column 0's
    majority word -- and centroid 1 stays (0, D, 0, ..).  Next pass: row0
    (16, 24) row1 (24, 16) row2 (40, 0) -> clusters [0, 1, 1], a fixed point
    (centroid 0 = (A, .., A) from row 0 alone, centroid 1 = (0, D, ..) from
    the bits both its rows share).  Without the chain column 32 stays 0 and
    the clusters stay [0, 0, 1].  Final masks at threshold 1.0: cluster 0 =
    row 0 alone, every bit invariant (mask all ones, bitval row 0); cluster 1
    = rows 1, 2, which differ exactly in A's bits of column 0 (mask
    0xFFFF0000 there, all ones elsewhere; bitval D in column 1)."""
    masks = np.full((2, 33), 0xFFFFFFFF, dtype=np.uint32)
    masks[1, 0] = 0xFFFF0000
    vals = np.zeros((2, 33), dtype=np.uint32)
    vals[0, 0] = 0x0000FFFF
    vals[1, 1] = 0xFF000000
    return masks, vals, np.array([0, 1, 1], dtype=np.int32)


def test_kmeans_or_chain_known_answer(oracle):
    """V > 32: the never-cleared accumulator of construct_bitmasks moves row 1
    to cluster 1 (an accumulator reset per column, or another stride, leaves it
    in cluster 0)."""
    assert list(oracle.random_shuffle(3, seed=1)) == [0, 2, 1]
    m, b, cl = oracle.kmeans(or_chain_data(), 2, seed=1)
    em, eb, ecl = or_chain_expected()
    assert np.array_equal(cl, ecl)
    assert np.array_equal(m.view(np.uint32), em) and np.array_equal(b.view(np.uint32), eb)


def test_oracle_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    """The oracle, built with AddressSanitizer and UndefinedBehaviorSanitizer
    (host code), run over random tables of every width and shape class by a C
    driver that also checks its round trips: no report, every case passes."""
    import os
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "oracle_san"
    r = subprocess.run(["gcc", "-O1", "-g", "-std=c11", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-fno-omit-frame-pointer", "-fopenmp", "-I" + os.path.join(root, "oracle"),
                        os.path.join(root, "oracle", "ibp_oracle.c"),
                        os.path.join(root, "tests", "oracle_sanitize_driver.c"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    for seed in (1, 2):
        r = subprocess.run([str(exe), str(seed), "120"], capture_output=True, text=True, timeout=240,
                           env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
        assert r.returncode == 0 and "passed" in r.stdout, r.stdout + r.stderr
