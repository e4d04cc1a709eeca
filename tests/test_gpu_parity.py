"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle,
bit-exact, on seeded inputs; plus size-independent properties at full size."""
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")


@pytest.fixture(scope="module")
def ibp():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import invariantbitpacking_amd as m
    m.index_errors(reset=True)
    return m


def _np(t):
    return t.detach().cpu().numpy()


def _u(a):
    return a.view(np.dtype("u%d" % a.itemsize))


def float_rows(n, v, w, seed, zero_frac=0.3, sd=0.1):
    """GNN-like rows: floats with exact zeros, viewed as w-byte integers."""
    rng = np.random.default_rng(seed)
    f = rng.normal(0, sd, size=(n, v * w // 4 if w >= 4 else v)).astype(np.float32)
    f[rng.random(f.shape) < zero_frac] = 0.0
    raw = f.view(np.uint8).reshape(n, -1)[:, : v * w]
    return np.ascontiguousarray(raw).view(np.dtype("i%d" % w))


def place(arr, where):
    t = torch.from_numpy(np.ascontiguousarray(arr))
    return t.cuda() if where == "device" else t.pin_memory()


# ------------------------------------------------------------------ preprocess

@pytest.mark.parametrize("w", [1, 2, 4, 8])
@pytest.mark.parametrize("threshold", [None, 0.9])
@pytest.mark.parametrize("shape", [(1000, 64), (257, 9), (300, 130), (64, 1)])
def test_preprocess_matches_oracle(ibp, oracle, w, threshold, shape):
    n, v = shape
    data = float_rows(n, v, w, seed=n + v + w)
    for where in ("device", "pinned"):
        print("preprocess", w, threshold, shape, where, flush=True)
        m, b = ibp.preprocess(place(data, where), threshold)
        om, ob, _, _ = oracle.preprocess(data, threshold)
        assert np.array_equal(_u(_np(m)), om) and np.array_equal(_u(_np(b)), ob)
        assert m.dtype == {1: torch.int8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[w]


def test_preprocess_golden(ibp, oracle):
    with open(GOLDEN) as f:
        cases = json.load(f)["preprocess"]
    for c in cases:
        rows = np.array([[int(x, 16) for x in r] for r in c["rows"]], dtype=np.uint8).view(np.int8)
        m, b = ibp.preprocess(place(rows, "device"), c["threshold"])
        assert [hex(x) for x in _u(_np(m))] == [hex(int(x, 16)) for x in c["mask"]], c["name"]
        assert [hex(x) for x in _u(_np(b))] == [hex(int(x, 16)) for x in c["bitval"]], c["name"]


def banded_rows(n, v, w, seed):
    """Rows whose bit k of column j takes the column's majority value in a
    fraction of the rows drawn from [0.6, 1.0]: column-bit counts fall on both
    sides of every sweep threshold, so every threshold saves a different number
    of bits and the choice among them depends on each score."""
    rng = np.random.default_rng(seed)
    B = 8 * w
    frac = rng.uniform(0.6, 1.0, size=(v, B))
    majority = rng.random((v, B)) < 0.5
    words = np.zeros((n, v), dtype=np.uint64)
    for k in range(B):
        bit = (rng.random((n, v)) < frac[:, k]) == majority[:, k]
        words |= bit.astype(np.uint64) << np.uint64(k)
    return words.astype(np.dtype("u%d" % w)).view(np.dtype("i%d" % w))


@pytest.mark.parametrize("threshold", [None, 0.9, 0.3])
@pytest.mark.parametrize("shape", [(3000, 64, 8), (2000, 50, 8), (1500, 48, 4), (777, 128, 4), (500, 130, 8),
                                   (4096, 2, 8), (100, 4, 4), (900, 66, 4)])
def test_preprocess_choice_matches_oracle(ibp, oracle, threshold, shape):
    """The chosen sweep step (C ABI output) and the masks, on data whose
    thresholds all score differently; covers the row-per-lane score path (rows
    of whole 16-byte chunks) and the lane-per-column one (900 x 66 x 4 B)."""
    from invariantbitpacking_amd import _capi
    n, v, w = shape
    thr = -1.0 if threshold is None else threshold
    for data in (banded_rows(n, v, w, seed=n + v), float_rows(n, v, w, seed=v + w)):
        print("preprocess choice", shape, threshold, flush=True)
        t = place(data, "device")
        mask = torch.zeros(v, dtype=t.dtype, device="cuda")
        bitval = torch.zeros(v, dtype=t.dtype, device="cuda")
        step = torch.full((1,), -7, dtype=torch.int32, device="cuda")
        ws_bytes = _capi.lib().ibp_preprocess_workspace_size(w, v, thr)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
        _capi.call("ibp_preprocess", w, t.data_ptr(), n, v, thr, mask.data_ptr(), bitval.data_ptr(), step.data_ptr(),
                   ws.data_ptr(), ws_bytes, torch.cuda.current_stream().cuda_stream)
        om, ob, ostep, _ = oracle.preprocess(data, threshold)
        assert int(step[0]) == ostep
        assert np.array_equal(_u(_np(mask)), om) and np.array_equal(_u(_np(bitval)), ob)


# ------------------------------------------------------------------ size + pack

SHAPES = [(200, 1), (200, 3), (500, 50), (700, 64), (300, 65), (200, 130), (64, 1024), (40, 2752)]


@pytest.mark.parametrize("w", [4, 8])
@pytest.mark.parametrize("shape", SHAPES)
def test_size_and_compress_match_oracle(ibp, oracle, w, shape):
    n, v = shape
    data = float_rows(n, v, w, seed=7 * n + v + w)
    om, ob, _, _ = oracle.preprocess(data, None)
    mask = torch.from_numpy(om.view(np.dtype("i%d" % w))).cuda()
    bitval = torch.from_numpy(ob.view(np.dtype("i%d" % w))).cuda()
    osz, _ = oracle.get_compress_size(data, om, ob)
    ocomp, obm = oracle.compress(data, om, ob)
    for where in ("device", "pinned"):
        print("compress", w, shape, where, flush=True)
        t = place(data, where)
        sz = ibp.get_compress_size(t, mask, bitval)
        torch.cuda.synchronize()
        assert np.array_equal(_np(sz), osz)
        bm = ibp.compress_inplace(t, mask, bitval)
        torch.cuda.synchronize()
        assert np.array_equal(_np(bm), obm), "bitmask"
        assert np.array_equal(_u(_np(t)), _u(ocomp)), "slots"


@pytest.mark.parametrize("w", [1, 2, 4, 8])
@pytest.mark.parametrize("v", [65, 130, 300, 1024])
def test_wide_rows_mixing_raw_and_packed_match_oracle(ibp, oracle, w, v):
    """Wide rows are packed into the slot buffer before their size is known
    and dropped when they come out raw.  Rows that pack and rows of random
    bits (raw) interleave here, so every raw row's speculative pack runs
    between packed rows of the same and of neighbouring waves: slots,
    bitmask, condensed stream and sizes must equal the oracle's for the
    in-place pack, the indexed pack, the condensed stream and clusters."""
    n = 480
    rng = np.random.default_rng(1000 + v + w)
    data, om, ob = masked_rows(n, v, w, seed=v * 10 + w, kind="runs2", flag_frac=0.97)
    raw = rng.random(n) < 0.4
    noise = rng.integers(0, 256, size=(n, v * w), dtype=np.uint8).view(data.dtype)
    data = np.ascontiguousarray(np.where(raw[:, None], noise, data))
    mask = torch.from_numpy(om.view(data.dtype)).cuda()
    bitval = torch.from_numpy(ob.view(data.dtype)).cuda()
    osz, _ = oracle.get_compress_size(data, om, ob)
    assert (osz == v * w).any() and (osz < v * w).any()  # both kinds present
    print("wide mixed", w, v, flush=True)
    t = place(data, "device")
    assert np.array_equal(_np(ibp.get_compress_size(t, mask, bitval)), osz)
    ocomp, obm = oracle.compress(data, om, ob)
    bm = ibp.compress_inplace(t, mask, bitval)
    torch.cuda.synchronize()
    assert np.array_equal(_np(bm), obm), "bitmask"
    assert np.array_equal(_u(_np(t)), _u(ocomp)), "slots"
    sel = rng.integers(0, n, size=300)
    ocomp_i, obm_i = oracle.compress(data, om, ob, index=sel)
    t = place(data, "device")
    bm = ibp.compress_inplace(t, mask, bitval, torch.from_numpy(sel).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(_np(bm), obm_i), "bitmask (index)"
    assert np.array_equal(_u(_np(t)[:len(sel)]), _u(ocomp_i)), "slots (index)"
    ws, wo = oracle.compress_condensed(data, om, ob, index=sel)
    stream, off = ibp.compress_condensed(place(data, "device"), mask, bitval, torch.from_numpy(sel).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(_np(off), wo) and np.array_equal(_u(_np(stream)), _u(ws)), "condensed"
    # two mask rows, the second never matching: rows of cluster 1 are raw
    masks2 = np.stack([om, om]).view(data.dtype)
    bitvals2 = np.stack([ob, ~ob]).view(data.dtype)
    cl = rng.integers(0, 2, size=n).astype(np.int32)
    oc, obmc = oracle.compress(data, masks2, bitvals2, clusters=cl)
    t = place(data, "device")
    bm = ibp.compress_inplace(t, torch.from_numpy(masks2).cuda(), torch.from_numpy(bitvals2).cuda(),
                              clusters=torch.from_numpy(cl).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(_np(bm), obmc) and np.array_equal(_u(_np(t)), _u(oc)), "clusters"


def masked_rows(n, v, w, seed, kind, flag_frac=0.8):
    """Rows drawn against a chosen mask/bitval so that about flag_frac of the
    elements are flagged.  kind picks the free bits of each column: 'runsR'
    (one to R runs, R <= 4), 'many' (alternating bits: more runs
    than a run table holds), 'none' (mask all ones: a flagged element carries
    no bits), 'mixed' (columns of the last three) or 'raw' (nothing flags)."""
    rng = np.random.default_rng(seed)
    B = 8 * w
    full = (1 << B) - 1
    masks = []
    for j in range(v):
        k = kind if kind != "mixed" else ("runs4", "many", "none")[j % 3]
        if k.startswith("runs"):
            # 1 + j % R runs (R = 1..4) in disjoint quarters of the word
            m = full
            for q in range(1 + j % int(k[4])):
                lo = int(rng.integers(1, B // 8 + 1))
                m &= ~(((1 << lo) - 1) << (q * (B // 4) + int(rng.integers(0, B // 4 - lo + 1))))
        elif k == "many":
            m = int(("01" * (B // 2))[: B], 2) if j % 2 else int(("0011" * (B // 4))[: B], 2)
        elif k == "none":
            m = full
        else:
            m = full
        masks.append(m)
    m = np.array(masks, dtype=np.uint64)
    b = np.array([int(rng.integers(0, 1 << 62)) for _ in range(v)], dtype=np.uint64) & m
    x = rng.integers(0, 1 << 63, size=(n, v), dtype=np.uint64) ^ rng.integers(0, 2, size=(n, v), dtype=np.uint64)
    if kind != "raw":
        flag = rng.random((n, v)) < flag_frac
        x = np.where(flag, (x & ~m) | b, x)
    dt = np.dtype("u%d" % w)
    return (x.astype(dt).view(np.dtype("i%d" % w)), m.astype(dt), b.astype(dt))


# Rows of whole 16-byte chunks (not 64 bytes modulo 128) take the row-per-lane
# tile size and pack kernels; the shapes cover one and an odd number of
# chunks, one- and two-word headers, short last tiles, and rows of a multiple
# of 128 bytes, whose chunks are permuted in LDS (128, 256, 384, 512 bytes).
TILE_SHAPES = [(500, 50, 8), (300, 2, 8), (1000, 62, 8), (777, 12, 4), (2000, 60, 4), (130, 36, 4), (64, 4, 4),
               (70000, 50, 8), (600, 64, 8), (300, 16, 8), (513, 48, 8), (700, 64, 4), (200, 32, 4)]


@pytest.mark.parametrize("shape", TILE_SHAPES)
@pytest.mark.parametrize("kind", ["float", "runs1", "runs2", "runs3", "runs4", "many", "none", "mixed", "raw"])
def test_tile_size_and_pack_match_oracle(ibp, oracle, shape, kind):
    n, v, w = shape
    if kind == "float":
        data = float_rows(n, v, w, seed=n + v + w)
        om, ob, _, _ = oracle.preprocess(data, None)
    else:
        data, om, ob = masked_rows(n, v, w, seed=n + 3 * v + w, kind=kind)
    mask = torch.from_numpy(om.view(np.dtype("i%d" % w))).cuda()
    bitval = torch.from_numpy(ob.view(np.dtype("i%d" % w))).cuda()
    ocomp, obm = oracle.compress(data, om, ob)
    osz, _ = oracle.get_compress_size(data, om, ob)
    print("tile pack", shape, kind, flush=True)
    t = place(data, "device")
    ctr = torch.tensor([5], dtype=torch.int64, device="cuda")
    sz = ibp.get_compress_size(t, mask, bitval, None, ctr)
    torch.cuda.synchronize()
    assert np.array_equal(_np(sz), osz), "sizes"
    assert int(ctr[0]) == 5 + int(osz.sum())
    bm = ibp.compress_inplace(t, mask, bitval)
    torch.cuda.synchronize()
    assert np.array_equal(_np(bm), obm), "bitmask"
    assert np.array_equal(_u(_np(t)), _u(ocomp)), "slots"
    out = ibp.decompress_fetch(t, mask, bitval, bm, torch.device("cuda"))
    assert np.array_equal(_np(out), data)


def test_golden_slots_on_gpu(ibp):
    with open(GOLDEN) as f:
        cases = json.load(f)["slots"]
    for c in cases:
        w = c["w"]
        dt = np.dtype("u%d" % w)
        row = np.array([int(x, 16) for x in c["row"]], dtype=dt)[None, :].view(np.dtype("i%d" % w))
        mask = torch.from_numpy(np.array([int(x, 16) for x in c["mask"]], dtype=dt).view(row.dtype)).cuda()
        bitval = torch.from_numpy(np.array([int(x, 16) for x in c["bitval"]], dtype=dt).view(row.dtype)).cuda()
        t = place(row, "pinned")
        sizes = ibp.get_compress_size(t, mask, bitval)
        torch.cuda.synchronize()  # asynchronous, like the reference
        assert int(sizes[0]) == c["size"], c["name"]
        bm = ibp.compress_inplace(t, mask, bitval)
        torch.cuda.synchronize()
        assert [hex(x) for x in _u(_np(t))[0]] == [hex(int(x, 16)) for x in c["slot"]], c["name"]
        out = ibp.decompress_fetch(t, mask, bitval, bm, torch.device("cuda"))
        assert np.array_equal(_np(out), row), c["name"]


@pytest.mark.parametrize("w,v", [(8, 8), (8, 3), (4, 5), (2, 17), (1, 31)])
def test_size_narrow_rows_with_index_and_counter(ibp, oracle, w, v):
    """Rows of fewer than 32 elements (several rows per wave step), whole
    tables and indices with repeats and out-of-range entries, against the
    oracle; the counter adds the sizes of the valid rows."""
    data = float_rows(1000, v, w, seed=v + w)
    om, ob, _, _ = oracle.preprocess(data, 0.9)
    dt = np.dtype("i%d" % w)
    mask, bitval = (torch.from_numpy(a.view(dt)).cuda() for a in (om, ob))
    full, _ = oracle.get_compress_size(data, om, ob)
    idx = np.array([5, 0, 999, 5, 12, 77, 3, 998, 500], dtype=np.int64)
    osz, _ = oracle.get_compress_size(data, om, ob, idx)
    for where in ("device", "pinned"):
        t = place(data, where)
        whole = ibp.get_compress_size(t, mask, bitval)
        torch.cuda.synchronize()  # asynchronous, as in the reference (pinned output for a pinned table)
        assert np.array_equal(_np(whole), full)
        ctr = torch.tensor([100], dtype=torch.int64, device="cuda")
        sz = ibp.get_compress_size(t, mask, bitval, torch.from_numpy(idx).cuda(), ctr)
        torch.cuda.synchronize()
        assert np.array_equal(_np(sz), osz)
        assert int(ctr[0]) == 100 + int(osz.sum())
    bad = torch.tensor([1, -1, 1000, 2], dtype=torch.int64, device="cuda")
    sz = ibp.get_compress_size(place(data, "device"), mask, bitval, bad)
    torch.cuda.synchronize()
    assert _np(sz)[1] == 0 and _np(sz)[2] == 0 and _np(sz)[0] == full[1] and _np(sz)[3] == full[2]


def test_size_with_index_and_counter(ibp, oracle):
    data = float_rows(400, 64, 8, seed=3)
    om, ob, _, _ = oracle.preprocess(data, 0.9)
    mask, bitval = (torch.from_numpy(a.view(np.int64)).cuda() for a in (om, ob))
    idx = np.array([5, 0, 399, 5, 12, 77], dtype=np.int64)
    osz, _ = oracle.get_compress_size(data, om, ob, idx)
    for where in ("device", "pinned"):
        t = place(data, where)
        for itype in (torch.int32, torch.int64):
            ctr = torch.tensor([100], dtype=torch.int64, device="cuda")
            sz = ibp.get_compress_size(t, mask, bitval, torch.from_numpy(idx).to(itype).cuda(), ctr)
            torch.cuda.synchronize()
            assert np.array_equal(_np(sz), osz)
            assert int(ctr[0]) == 100 + int(osz.sum())
        pinned_ctr = torch.tensor([7], dtype=torch.int64).pin_memory()
        ibp.get_compress_size(t, mask, bitval, None, pinned_ctr)
        full, _ = oracle.get_compress_size(data, om, ob)
        assert int(pinned_ctr[0]) == 7 + int(full.sum())
    L = ibp.get_single_compress_len(place(data, "device"), mask, bitval)
    assert L == int(int(full.sum()) / data.nbytes * 64)


def test_compress_with_index(ibp, oracle):
    data = float_rows(128, 64, 8, seed=9)
    om, ob, _, _ = oracle.preprocess(data, None)
    mask, bitval = (torch.from_numpy(a.view(np.int64)).cuda() for a in (om, ob))
    idx = np.array([3, 3, 100, 1, 0, 64, 127], dtype=np.int64)
    ocomp, obm = oracle.compress(data, om, ob, idx)
    t = place(data, "pinned")
    bm = ibp.compress_inplace(t, mask, bitval, torch.from_numpy(idx).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(_np(bm), obm)
    assert np.array_equal(_u(_np(t)[: len(idx)]), _u(ocomp))


# ------------------------------------------------------------------ fetch

def _compressed(ibp, oracle, n, v, w, seed, where, threshold=None):
    data = float_rows(n, v, w, seed=seed)
    om, ob, _, _ = oracle.preprocess(data, threshold)
    mask = torch.from_numpy(om.view(np.dtype("i%d" % w))).cuda()
    bitval = torch.from_numpy(ob.view(np.dtype("i%d" % w))).cuda()
    ocomp, obm = oracle.compress(data, om, ob)
    comp = place(ocomp.view(np.dtype("i%d" % w)), where)
    bitmask = torch.from_numpy(obm).cuda()
    return data, comp, mask, bitval, bitmask, (om, ob, ocomp, obm)


@pytest.mark.parametrize("w", [4, 8])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("where", ["device", "pinned"])
def test_fetch_matches_oracle(ibp, oracle, w, shape, where):
    n, v = shape
    data, comp, mask, bitval, bitmask, (om, ob, ocomp, obm) = _compressed(ibp, oracle, n, v, w, 11 * n + v, where)
    rng = np.random.default_rng(v)
    idx = rng.integers(0, n, size=3 * n + 5)  # duplicates, unsorted
    want = oracle.decompress_fetch(ocomp, om, ob, obm, idx)
    assert np.array_equal(_u(want), _u(data[idx]))
    L = ibp.get_single_compress_len(comp, mask, bitval)
    for itype in (torch.int64, torch.int32):
        for comp_len in (None, L, 1, v):
            print("fetch", w, shape, where, itype, comp_len, flush=True)
            out = ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), comp_len,
                                       torch.from_numpy(idx).to(itype).cuda())
            torch.cuda.synchronize()
            assert np.array_equal(_u(_np(out)), _u(want))
    out = ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"))
    assert np.array_equal(_np(out), data)


@pytest.mark.parametrize("imp", [-1, 0, 1, 5])
def test_fetch_impls_identical(ibp, oracle, imp):
    data, comp, mask, bitval, bitmask, _ = _compressed(ibp, oracle, 333, 64, 8, 5, "pinned")
    idx = torch.randint(0, 333, (1000,), generator=torch.Generator().manual_seed(imp + 2)).cuda()
    out = ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), None, idx, 320, 512, imp)
    assert np.array_equal(_np(out), data[_np(idx)])


def test_fetch_raw_gather_impl2(ibp, oracle):
    data, comp, mask, bitval, bitmask, (om, ob, ocomp, obm) = _compressed(ibp, oracle, 200, 64, 8, 6, "pinned")
    idx = torch.arange(199, -1, -1).cuda()
    out = ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), None, idx, None, None, 2)
    assert np.array_equal(_u(_np(out)), _u(ocomp[_np(idx)]))


def test_fetch_output_tensor_and_host_output(ibp, oracle):
    data, comp, mask, bitval, bitmask, _ = _compressed(ibp, oracle, 500, 50, 8, 8, "device")
    idx = torch.randint(0, 500, (777,), generator=torch.Generator().manual_seed(3))
    out = torch.full((777, 50), -1, dtype=torch.int64, device="cuda")
    r = ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), None, idx.cuda(), None, None, None,
                             out)
    assert r.data_ptr() == out.data_ptr()
    assert np.array_equal(_np(out), data[idx.numpy()])
    host = ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cpu"), None, idx.cuda())
    torch.cuda.synchronize()
    assert host.device.type == "cpu" and host.is_pinned()
    assert np.array_equal(host.numpy(), data[idx.numpy()])
    with pytest.raises(RuntimeError, match="shape"):
        ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), None, idx.cuda(), None, None, None,
                             torch.empty((776, 50), dtype=torch.int64, device="cuda"))
    with pytest.raises(RuntimeError, match="Index array"):
        ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), None, idx)


def test_fetch_edge_cases(ibp, oracle):
    data, comp, mask, bitval, bitmask, _ = _compressed(ibp, oracle, 100, 64, 8, 10, "pinned")
    empty = ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), None,
                                 torch.empty(0, dtype=torch.int64, device="cuda"))
    assert tuple(empty.shape) == (0, 64)
    ibp.index_errors(reset=True)
    idx = torch.tensor([0, 100, -1, 99, 1 << 40], dtype=torch.int64, device="cuda")
    out = _np(ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), None, idx))
    assert np.array_equal(out[0], data[0]) and np.array_equal(out[3], data[99])
    assert not out[1].any() and not out[2].any() and not out[4].any()
    assert ibp.index_errors(reset=True) == 3
    # single row, 1-D dataset
    one = torch.from_numpy(data[7].copy()).pin_memory()
    bm1 = ibp.compress_inplace(one, mask, bitval)
    out1 = ibp.decompress_fetch(one, mask, bitval, bm1, torch.device("cuda"))
    assert tuple(out1.shape) == (1, 64) and np.array_equal(_np(out1)[0], data[7])


@pytest.mark.parametrize("w", [4, 8])
def test_empty_tables(ibp, oracle, w):
    """A table of zero rows (on the GPU: an empty host tensor is not pinned,
    so the calls reject it, as the reference's checks do) through every call:
    zero masks (nothing to save), empty sizes, bitmask, stream and offsets,
    fetches of zero rows; k-means gives the oracle's all-zero centroids and no
    clusters."""
    v = 8
    x = np.zeros((0, v), dtype=np.dtype("i%d" % w))
    om, ob, _, _ = oracle.preprocess(x, 0.9)
    okm, okb, okc = oracle.kmeans(x, 4, seed=1)
    for where in ("device",):
        print("empty table", w, where, flush=True)
        t = place(x, where)
        for thr in (0.9, None):
            m, b = ibp.preprocess(t, thr)
            assert np.array_equal(_u(_np(m)), om) and np.array_equal(_u(_np(b)), ob)
        m, b = ibp.preprocess(place(np.ones((3, v), dtype=x.dtype), "device"), 0.9)
        sz = ibp.get_compress_size(t, m, b)
        assert tuple(sz.shape) == (0,)
        bm = ibp.compress_inplace(t, m, b)
        assert tuple(bm.shape) == (0,)
        stream, off = ibp.compress_condensed(t, m, b)
        assert stream.numel() == 0 and off.numel() == 0
        out = ibp.decompress_fetch(t, m, b, bm, torch.device("cuda"))
        assert tuple(out.shape) == (0, v)
        back = ibp.decompress(stream, off, m, b, torch.device("cuda"), v)
        assert tuple(back.shape) == (0, v)
        km, kb, kc = ibp.preprocess_kmeans(t, 4, 0.85, seed=1)
        assert np.array_equal(_u(_np(km)), okm) and np.array_equal(_u(_np(kb)), okb) and kc.numel() == 0


@pytest.mark.parametrize("where,v", [("device", 50), ("pinned", 50), ("device", 130)])
def test_fetch_captured_in_a_torch_graph(ibp, oracle, where, v):
    """The plain decompress_fetch with an output tensor allocates nothing and
    never synchronises, so a torch CUDA graph can capture it: each replay,
    with the index refilled and eager work and synchronisations between
    replays, equals the eager fetch and the original rows."""
    n, w = 3000, 8
    data, comp, mask, bitval, bitmask, _ = _compressed(ibp, oracle, n, v, w, 77 + v, where)
    dev = torch.device("cuda")
    rng = np.random.default_rng(v)
    idx = torch.from_numpy(rng.integers(0, n, size=1024)).cuda()
    out = torch.empty((1024, v), dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # first calls (kernel loading) outside the capture
        ibp.decompress_fetch(comp, mask, bitval, bitmask, dev, None, idx, None, None, None, out)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ibp.decompress_fetch(comp, mask, bitval, bitmask, dev, None, idx, None, None, None, out)
    for rep in range(4):
        print("graph replay", where, v, rep, flush=True)
        sel = rng.integers(0, n, size=1024)
        idx.copy_(torch.from_numpy(sel))
        out.fill_(-1)
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(_np(out), data[sel]), rep
        eager = ibp.decompress_fetch(comp, mask, bitval, bitmask, dev, None, idx)
        torch.cuda.synchronize()
        assert torch.equal(eager, out), rep


def test_fetch_full_width_fields(ibp, oracle):
    """Raw 64-bit elements at 64-aligned stream positions and flagged elements
    with mask 0 (the full-width shift cases) through compress + fetch."""
    rng = np.random.default_rng(0)
    v = 64
    mask = np.zeros(v, dtype=np.uint64)
    mask[1::2] = np.uint64(0xFFFFFFFFFFFFFFF0)
    data = rng.integers(0, 1 << 62, size=(256, v), dtype=np.uint64)
    data[:, 1::2] &= np.uint64(0xF)  # flagged under mask, 4 free bits each
    data[::3, 1::4] |= np.uint64(1 << 63)  # some unflagged -> raw 64-bit fields
    bitval = np.zeros(v, dtype=np.uint64)
    ocomp, obm = oracle.compress(data, mask, bitval)
    assert obm.any()
    m = torch.from_numpy(mask.view(np.int64)).cuda()
    b = torch.from_numpy(bitval.view(np.int64)).cuda()
    t = place(data.view(np.int64), "pinned")
    bm = ibp.compress_inplace(t, m, b)
    torch.cuda.synchronize()
    assert np.array_equal(_u(_np(t)), ocomp) and np.array_equal(_np(bm), obm)
    out = ibp.decompress_fetch(t, m, b, bm, torch.device("cuda"))
    assert np.array_equal(_u(_np(out)), data)


# ------------------------------------------------------------------ kmeans

@pytest.mark.parametrize("w,n,v,k", [(4, 60, 40, 4), (8, 60, 70, 3), (8, 300, 64, 16), (4, 10, 5, 16),
                                     (8, 4096, 130, 16), (4, 3000, 300, 20), (8, 6000, 64, 4),
                                     (4, 1100000, 2, 3), (8, 300000, 20, 12), (8, 2048, 520, 17),
                                     (4, 1, 1, 1), (8, 1, 3, 2), (8, 65, 17, 16), (4, 129, 33, 5),
                                     (8, 1200000, 8, 16), (4, 5000, 2, 4100)])
def test_kmeans_matches_oracle(ibp, oracle, w, n, v, k):
    """Small shapes; 4096 x 130 (one block per cluster and column tile; the
    row-per-lane assignment over nine column groups, the last one two words
    wide, summed by the argmin kernel); K = 20 (the wave-per-row-group
    assignment, two groups of 16 centroids); 6000 rows in 4 clusters; 1.1M
    and 1.2M rows (beyond the one-block sort: histogram, chunk list and
    scatter over many blocks); 300,000 8-byte rows (one column group: the
    assignment kernel picks the nearest centroid itself); 8-byte rows with
    K = 17 (the wave-per-row-group assignment, a second group of one
    centroid); K = 4100 (beyond the LDS histograms: per-row atomics and a
    bit count per cluster into zeroed counters); one row, K > N, partial row
    tiles and odd widths."""
    rng = np.random.default_rng(n + v + k)
    dt = np.dtype("u%d" % w)
    base = rng.integers(0, 1 << 62, size=(4, v), dtype=np.uint64).astype(dt)
    data = base[rng.integers(0, 4, size=n)] ^ rng.integers(0, 16, size=(n, v), dtype=np.uint64).astype(dt)
    data = data.view(np.dtype("i%d" % w))
    om, ob, ocl = oracle.kmeans(data, k, seed=3)
    for where in ("device", "pinned"):
        print("kmeans", w, n, v, k, where, flush=True)
        m, b, cl = ibp.preprocess_kmeans(place(data, where), k, 0.85, seed=3)
        assert cl.is_pinned() and cl.dtype == torch.int32
        assert np.array_equal(cl.numpy(), ocl)
        assert np.array_equal(_u(_np(m)), om) and np.array_equal(_u(_np(b)), ob)


def test_kmeans_or_chain_known_answer_on_gpu(ibp):
    """The hand-derived stride-32 OR-chain case (33 columns, K = 2): the
    never-cleared accumulator carries column 0's majority word into column 32
    and moves row 1 to cluster 1."""
    from test_oracle import or_chain_data, or_chain_expected
    em, eb, ecl = or_chain_expected()
    for where in ("device", "pinned"):
        m, b, cl = ibp.preprocess_kmeans(place(or_chain_data(), where), 2, seed=1)
        assert np.array_equal(cl.numpy(), ecl)
        assert np.array_equal(_np(m).view(np.uint32), em) and np.array_equal(_np(b).view(np.uint32), eb)


def test_kmeans_mask_row0_used_by_compress(ibp, oracle):
    """A [K, V] mask is used through its first row, as in the reference."""
    data = float_rows(64, 16, 8, seed=1)
    m, b, _ = ibp.preprocess_kmeans(place(data, "device"), 4, seed=1)
    om, ob = _u(_np(m))[0], _u(_np(b))[0]
    ocomp, obm = oracle.compress(data, om, ob)
    t = place(data, "pinned")
    bm = ibp.compress_inplace(t, m, b)
    torch.cuda.synchronize()
    assert np.array_equal(_np(bm), obm) and np.array_equal(_u(_np(t)), _u(ocomp))


# ------------------------------------------------------------------ full size

def test_gnn_features_full_size_round_trip(ibp, oracle):
    """65536 x 64 int64 (fp32 N(0,0.1), 30% zeros) at threshold 0.9, pinned,
    4096 random indices: the fetched rows equal the original rows."""
    from invariantbitpacking_amd import workloads
    data, idx = workloads.gnn_features()
    x = data.numpy()
    pinned = data.clone().pin_memory()
    mask, bitval = ibp.preprocess(pinned, 0.9)
    om, ob, _, _ = oracle.preprocess(x, 0.9)
    assert np.array_equal(_u(_np(mask)), om) and np.array_equal(_u(_np(bitval)), ob)
    sizes = ibp.get_compress_size(pinned, mask, bitval)
    bm = ibp.compress_inplace(pinned, mask, bitval)
    torch.cuda.synchronize()
    s = sizes.numpy()
    assert np.all(s <= 512) and np.all(s % 8 == 0)
    assert np.array_equal(np.unpackbits(_np(bm).view(np.uint8), bitorder="little")[: len(s)].astype(bool), s < 512)
    L = ibp.get_single_compress_len(data.clone().pin_memory(), mask, bitval)
    out = ibp.decompress_fetch(pinned, mask, bitval, bm, torch.device("cuda"), L, idx.cuda())
    assert torch.equal(out.cpu(), data[idx])
    dev = pinned.cuda()
    out2 = ibp.decompress_fetch(dev, mask, bitval, bm, torch.device("cuda"), L, idx.cuda())
    assert torch.equal(out2.cpu(), data[idx])
