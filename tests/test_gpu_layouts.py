"""GPU parity of the condensed stream (compress_condensed / compress /
decompress), per-cluster masks (clusters=) and 1/2-byte elements, bit-exact
against the oracle on the same inputs."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ibp():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import invariantbitpacking_amd as m
    return m


def rows_of(n, v, w, seed, zero_frac=0.3):
    """fp32 N(0, 0.05) with zeros, viewed as v elements of w bytes."""
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 0.05, size=(n, max(1, v * w // 4) + 1)).astype(np.float32)
    f[rng.random(f.shape) < zero_frac] = 0.0
    b = f.view(np.uint8)[:, :v * w].copy()
    return b.view(np.dtype("i%d" % w))


def group_masks(oracle, x, k):
    """K distinct mask rows: preprocess over K groups of rows."""
    n = x.shape[0]
    ms, bs = [], []
    for c in range(k):
        m, b, _, _ = oracle.preprocess(x[c * n // k:(c + 1) * n // k], 0.7 + 0.1 * (c % 3))
        ms.append(m)
        bs.append(b)
    return np.stack(ms).view(x.dtype), np.stack(bs).view(x.dtype)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _pinned(a):
    return torch.from_numpy(np.ascontiguousarray(a)).pin_memory()


SHAPES = [(300, 1), (257, 7), (500, 50), (400, 64), (300, 65), (120, 300), (40, 1100)]


@pytest.mark.parametrize("w", [1, 2, 4, 8])
@pytest.mark.parametrize("n,v", SHAPES)
def test_condensed_matches_oracle(ibp, oracle, w, n, v):
    x = rows_of(n, v, w, seed=n + v + w)
    xd = _dev(x)
    mask, bitval = ibp.preprocess(xd, 0.8)
    m, b = mask.cpu().numpy(), bitval.cpu().numpy()
    want_stream, want_off = oracle.compress_condensed(x, m, b)
    for where in ("device", "pinned"):
        src = xd if where == "device" else _pinned(x)
        stream, off = ibp.compress_condensed(src, mask, bitval)
        torch.cuda.synchronize()
        assert stream.is_cuda == (where == "device") and off.is_cuda == (where == "device")
        assert np.array_equal(off.cpu().numpy(), want_off), where
        assert np.array_equal(stream.cpu().numpy().view(want_stream.dtype), want_stream), where
        # decompress: all rows, then random rows (int32 and int64 indices)
        got = ibp.decompress(stream, off, mask, bitval, torch.device("cuda"), v)
        torch.cuda.synchronize()
        assert torch.equal(got.cpu(), torch.from_numpy(x)), where
        idx = np.random.default_rng(1).integers(0, n, size=2 * n)
        for dt in (torch.int64, torch.int32):
            got = ibp.decompress(stream, off, mask, bitval, torch.device("cuda"), v,
                                 torch.from_numpy(idx).to(dt).cuda(), comp_len=max(1, v // 2))
            torch.cuda.synchronize()
            assert torch.equal(got.cpu(), torch.from_numpy(x[idx])), (where, dt)
    # ibp.compress is compress_condensed
    s2, o2 = ibp.compress(xd, mask, bitval)
    torch.cuda.synchronize()
    assert torch.equal(s2.cpu(), stream.cpu()) and torch.equal(o2.cpu(), off.cpu())


def test_condensed_with_index_and_output_tensor(ibp, oracle):
    n, v, w = 2000, 40, 8
    x = rows_of(n, v, w, seed=5)
    xd = _dev(x)
    mask, bitval = ibp.preprocess(xd, None)
    m, b = mask.cpu().numpy(), bitval.cpu().numpy()
    sel = np.random.default_rng(2).integers(0, n, size=777)
    ws, wo = oracle.compress_condensed(x, m, b, index=sel)
    for idx in (torch.from_numpy(sel).cuda(), torch.from_numpy(sel).to(torch.int32).pin_memory()):
        stream, off = ibp.compress_condensed(xd, mask, bitval, idx)
        torch.cuda.synchronize()
        assert np.array_equal(off.cpu().numpy(), wo)
        assert np.array_equal(stream.cpu().numpy().view(ws.dtype), ws)
    out = torch.full((777, v), 3, dtype=torch.int64, device="cuda")
    r = ibp.decompress(stream, off, mask, bitval, torch.device("cuda"), v, None, out)
    torch.cuda.synchronize()
    assert r is out and torch.equal(out.cpu(), torch.from_numpy(x[sel]))


def test_condensed_invalid_rows_are_zero_and_counted(ibp):
    n, v = 500, 50
    x = rows_of(n, v, 8, seed=9)
    xd = _dev(x)
    mask, bitval = ibp.preprocess(xd, 0.8)
    stream, off = ibp.compress_condensed(xd, mask, bitval)
    torch.cuda.synchronize()
    ibp.index_errors(reset=True)
    idx = torch.tensor([0, n, 7, -1, n - 1], dtype=torch.int64, device="cuda")
    got = ibp.decompress(stream, off, mask, bitval, torch.device("cuda"), v, idx).cpu()
    assert ibp.index_errors(reset=True) == 2
    assert torch.equal(got[0], torch.from_numpy(x[0])) and torch.equal(got[2], torch.from_numpy(x[7]))
    assert not got[1].any() and not got[3].any()
    # offsets that point outside the stream: zero rows, counted, nothing read out of bounds
    bad = off.clone()
    bad[3] = stream.numel() * 8 + 800
    got = ibp.decompress(stream, bad, mask, bitval, torch.device("cuda"), v).cpu()
    assert ibp.index_errors(reset=True) >= 1
    assert not got[3].any() and torch.equal(got[2], torch.from_numpy(x[2]))


def mixed_rows(oracle, n, v, w, seed, zero_frac=0.3):
    """GNN-like rows (they pack under preprocess at 0.8) with about 40% of the
    rows replaced by random bytes (they stay raw), and their mask/bitval."""
    rng = np.random.default_rng(seed)
    x = rows_of(n, v, w, seed=seed, zero_frac=zero_frac)
    m, b, _, _ = oracle.preprocess(x, 0.8)
    raw = rng.random(n) < 0.4
    noise = rng.integers(0, 256, size=(n, v * w), dtype=np.uint8).view(x.dtype)
    return np.ascontiguousarray(np.where(raw[:, None], noise, x)), m, b


@pytest.mark.parametrize("n,v,w,zf", [(3000, 50, 8, 0.3), (1000, 64, 8, 0.3), (777, 48, 8, 0.3), (2000, 60, 4, 0.3),
                                      (900, 64, 4, 0.3), (129, 16, 8, 0.9), (500, 32, 8, 0.9), (700, 36, 4, 0.9)])
def test_condensed_whole_table_mixed_rows(ibp, oracle, n, v, w, zf):
    """compress_condensed of a whole table with one mask (the row-per-lane
    tile path for rows of whole 16-byte chunks): packed and raw rows
    interleave, the stream and offsets equal the oracle's, device and pinned,
    and the stream decompresses to the rows."""
    x, m, b = mixed_rows(oracle, n, v, w, seed=n + v, zero_frac=zf)
    ws, wo = oracle.compress_condensed(x, m, b)
    sizes = np.diff(np.concatenate([[0], wo]))
    assert (sizes == v * w).any() and (sizes < v * w).any()
    mask, bitval = _dev(m.view(x.dtype)), _dev(b.view(x.dtype))
    for where in ("device", "pinned"):
        print("condensed whole table", n, v, w, where, flush=True)
        src = _dev(x) if where == "device" else _pinned(x)
        stream, off = ibp.compress_condensed(src, mask, bitval)
        torch.cuda.synchronize()
        assert np.array_equal(off.cpu().numpy(), wo), where
        assert np.array_equal(stream.cpu().numpy().view(ws.dtype), ws), where
        got = ibp.decompress(stream, off, mask, bitval, torch.device("cuda"), v)
        torch.cuda.synchronize()
        assert torch.equal(got.cpu(), torch.from_numpy(x)), where


def test_condensed_pack_with_wrong_offsets_writes_only_consistent_rows(ibp, oracle):
    """The pack half of compress_condensed through the C ABI with offsets whose
    row lengths are not the rows' sizes: those rows are counted and left
    unwritten, every other row is written exactly (whole-table tile path)."""
    import ctypes

    from invariantbitpacking_amd import _capi
    n, v, w = 1000, 50, 8
    x, m, b = mixed_rows(oracle, n, v, w, seed=77)
    ws, wo = oracle.compress_condensed(x, m, b)
    sizes = np.diff(np.concatenate([[0], wo]))
    # rows 5 and 6 change lengths (row 5 grows by a word, row 6 shrinks), and
    # one raw and one packed row elsewhere swap lengths with their neighbour
    bad = wo.copy()
    bad[5] += w
    r_raw = int(np.flatnonzero(sizes[20:] == v * w)[0]) + 20
    bad[r_raw - 1] += w  # row r_raw - 1 one word longer, row r_raw one word shorter
    wrong = {5, 6, r_raw - 1, r_raw}
    wrong = {i for i in wrong if (bad[i] - (bad[i - 1] if i else 0)) != sizes[i]}
    xd = _dev(x)
    mask, bitval = _dev(m.view(x.dtype)), _dev(b.view(x.dtype))
    offs = torch.from_numpy(bad).cuda()
    total = int(bad[-1])
    stream = torch.full((total // w,), 0x5A5A5A5A5A5A5A5A, dtype=torch.int64, device="cuda")
    ibp.index_errors(reset=True)
    lay = _capi.Layout(offs.data_ptr(), total, None, 1)
    _capi.call("ibp_compress_ex", w, xd.data_ptr(), stream.data_ptr(), n, v, mask.data_ptr(), bitval.data_ptr(),
               None, 0, n, None, ctypes.byref(lay), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert ibp.index_errors(reset=True) == len(wrong)
    got = stream.cpu().numpy().view(np.uint8)
    want = ws.view(np.uint8)
    ends = np.concatenate([[0], wo])
    bends = np.concatenate([[0], bad])
    for i in range(n):
        seg = got[bends[i]:bends[i + 1]]
        if i in wrong:
            assert (seg == 0x5A).all(), i
        else:
            assert np.array_equal(seg, want[ends[i]:ends[i + 1]]), i


@pytest.mark.parametrize("w", [4, 8])
@pytest.mark.parametrize("n,v", [(600, 50), (500, 64), (300, 130), (64, 1024)])
def test_clusters_match_oracle(ibp, oracle, w, n, v):
    x = rows_of(n, v, w, seed=11 * v + w)
    xd = _dev(x)
    masks, bitvals, clusters = ibp.preprocess_kmeans(xd, 4, None, seed=3)
    m, b, c = masks.cpu().numpy(), bitvals.cpu().numpy(), clusters.numpy()
    ws, _ = oracle.get_compress_size(x, m, b, clusters=c)
    sizes = ibp.get_compress_size(xd, masks, bitvals, clusters=clusters)
    torch.cuda.synchronize()
    assert np.array_equal(sizes.cpu().numpy(), ws)
    want_slots, want_bm = oracle.compress(x, m, b, clusters=c)
    for where in ("device", "pinned"):
        data = xd.clone() if where == "device" else _pinned(x)
        bm = ibp.compress_inplace(data, masks, bitvals, clusters=clusters)
        assert np.array_equal(bm.cpu().numpy(), want_bm), where
        assert np.array_equal(data.cpu().numpy().view(want_slots.dtype), want_slots), where
        out = ibp.decompress_fetch(data, masks, bitvals, bm, torch.device("cuda"), clusters=clusters.cuda())
        torch.cuda.synchronize()
        assert torch.equal(out.cpu(), torch.from_numpy(x)), where
        idx = np.random.default_rng(4).integers(0, n, size=999)
        out = ibp.decompress_fetch(data, masks, bitvals, bm, torch.device("cuda"), None,
                                   torch.from_numpy(idx).cuda(), clusters=clusters)
        torch.cuda.synchronize()
        assert torch.equal(out.cpu(), torch.from_numpy(x[idx])), where
    # condensed + clusters
    stream, off = ibp.compress_condensed(xd, masks, bitvals, clusters=clusters)
    torch.cuda.synchronize()
    s_want, o_want = oracle.compress_condensed(x, m, b, clusters=c)
    assert np.array_equal(off.cpu().numpy(), o_want)
    assert np.array_equal(stream.cpu().numpy().view(s_want.dtype), s_want)
    back = ibp.decompress(stream, off, masks, bitvals, torch.device("cuda"), v, clusters=clusters)
    torch.cuda.synchronize()
    assert torch.equal(back.cpu(), torch.from_numpy(x))


def test_clusters_index_uses_source_rows_and_bad_ids_are_counted(ibp, oracle):
    n, v, k = 800, 48, 5
    x = rows_of(n, v, 8, seed=21)
    masks_np, bitvals_np = group_masks(oracle, x, k)
    masks, bitvals = _dev(masks_np), _dev(bitvals_np)
    c = np.random.default_rng(6).integers(0, k, size=n).astype(np.int32)
    cl = _dev(c)
    sel = np.random.default_rng(7).integers(0, n, size=300)
    ws, wo = oracle.compress_condensed(x, masks_np, bitvals_np, index=sel, clusters=c)
    stream, off = ibp.compress_condensed(_dev(x), masks, bitvals, torch.from_numpy(sel).cuda(), clusters=cl)
    torch.cuda.synchronize()
    assert np.array_equal(off.cpu().numpy(), wo) and np.array_equal(stream.cpu().numpy().view(ws.dtype), ws)
    stored_cl = cl[torch.from_numpy(sel).cuda()]
    back = ibp.decompress(stream, off, masks, bitvals, torch.device("cuda"), v, clusters=stored_cl)
    torch.cuda.synchronize()
    assert torch.equal(back.cpu(), torch.from_numpy(x[sel]))
    # compress_inplace with an index: slot i receives row sel[i], coded with
    # that row's mask row; the slots past the index are untouched
    slots = _dev(x)
    bm = ibp.compress_inplace(slots, masks, bitvals, torch.from_numpy(sel).cuda(), clusters=cl)
    want, wbm = oracle.compress(x, masks_np, bitvals_np, index=sel, clusters=c)
    assert np.array_equal(bm.cpu().numpy(), wbm)
    assert np.array_equal(slots[:300].cpu().numpy().view(want.dtype), want)
    assert torch.equal(slots[300:].cpu(), torch.from_numpy(x[300:]))
    # an invalid cluster id gives a zero row and is counted
    bad = stored_cl.clone()
    bad[5] = k
    bad[6] = -1
    ibp.index_errors(reset=True)
    got = ibp.decompress(stream, off, masks, bitvals, torch.device("cuda"), v, clusters=bad).cpu()
    assert ibp.index_errors(reset=True) == 2
    assert not got[5].any() and not got[6].any() and torch.equal(got[7], torch.from_numpy(x[sel[7]]))


@pytest.mark.parametrize("w", [1, 2])
@pytest.mark.parametrize("n,v", [(400, 1), (300, 9), (500, 40), (400, 64), (300, 65), (100, 700)])
def test_narrow_elements_match_oracle(ibp, oracle, w, n, v):
    x = rows_of(n, v, w, seed=31 * v + w, zero_frac=0.5)
    xd = _dev(x)
    mask, bitval = ibp.preprocess(xd, 0.75)
    m, b = mask.cpu().numpy(), bitval.cpu().numpy()
    want, wbm = oracle.compress(x, m, b)
    for where in ("device", "pinned"):
        data = xd.clone() if where == "device" else _pinned(x)
        bm = ibp.compress_inplace(data, mask, bitval)
        assert np.array_equal(bm.cpu().numpy(), wbm), where
        assert np.array_equal(data.cpu().numpy().view(want.dtype), want), where
        idx = np.random.default_rng(8).integers(0, n, size=3 * n)
        out = ibp.decompress_fetch(data, mask, bitval, bm, torch.device("cuda"), None, torch.from_numpy(idx).cuda())
        torch.cuda.synchronize()
        assert torch.equal(out.cpu(), torch.from_numpy(x[idx])), where
        assert np.array_equal(out.cpu().numpy(), oracle.decompress_fetch(want, m, b, wbm, index=idx))


def test_offsets_scan_at_scale(ibp):
    """5M rows (1221 scan tiles, more than one block of tile sums): offsets
    equal the running sum of get_compress_size, and the stream round-trips."""
    n, v = 5_000_000, 16
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randint(0, 1 << 12, (n, v), dtype=torch.int32, device="cuda", generator=g)
    big = torch.rand((n, v), device="cuda", generator=g) < 0.2  # rows differ in how many elements pack
    x[big] = torch.randint(-2**31, 2**31 - 1, (int(big.sum()),), dtype=torch.int32, device="cuda", generator=g)
    x[torch.rand((n, 1), device="cuda", generator=g).expand(n, v) < 0.05] = -1  # rows that never pack
    mask, bitval = ibp.preprocess(x, 0.75)
    sizes = ibp.get_compress_size(x, mask, bitval)
    stream, off = ibp.compress_condensed(x, mask, bitval)
    torch.cuda.synchronize()
    assert torch.equal(off, torch.cumsum(sizes, 0))
    assert len(set(sizes[:100000].tolist())) > 2
    idx = torch.randint(0, n, (300000,), device="cuda", generator=g)
    got = ibp.decompress(stream, off, mask, bitval, torch.device("cuda"), v, idx)
    torch.cuda.synchronize()
    assert torch.equal(got, x[idx])
    whole = ibp.decompress(stream, off, mask, bitval, torch.device("cuda"), v)
    torch.cuda.synchronize()
    assert torch.equal(whole, x)
