"""ibp_condensed_offsets_noalloc (the allocation-free form of
ibp_condensed_offsets) against ibp_condensed_offsets, through the C ABI:
identical offsets on two calls in a row, the second smaller, with and without
an index and per-row clusters, and the stated limits."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import invariantbitpacking_amd as ibp  # noqa: F401  (loads the library)
    from invariantbitpacking_amd import _capi
    return _capi


def _rows(n, v, w, seed):
    """fp32 N(0, 0.05) with 30% zeros, generated on the GPU, viewed as v
    elements of w bytes."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    f = torch.randn((n, v * w // 4), generator=g, device="cuda") * 0.05
    f[torch.rand(f.shape, generator=g, device="cuda") < 0.3] = 0.0
    return f.view({4: torch.int32, 8: torch.int64}[w])


def _offsets(capi, name, x, mask, bitval, index=None, clusters=None, k=1):
    n = x.shape[0] if index is None else index.numel()
    out = torch.full((n,), -7, dtype=torch.int64, device="cuda")
    lay = None
    if clusters is not None:
        lay = ctypes.byref(capi.Layout(None, 0, clusters.data_ptr(), k))
    stream = torch.cuda.current_stream().cuda_stream
    capi.call(name, x.element_size(), x.data_ptr(), x.shape[0], x.shape[1], mask.data_ptr(), bitval.data_ptr(),
              None if index is None else index.data_ptr(), 0 if index is None else index.element_size(), n,
              out.data_ptr(), lay, stream)
    return out


def _masks(x, k):
    import invariantbitpacking_amd as ibp
    ms, bs = [], []
    for c in range(k):
        m, b = ibp.preprocess(x[c::k].contiguous(), 0.7 + 0.1 * (c % 3))
        ms.append(m)
        bs.append(b)
    return torch.stack(ms).contiguous(), torch.stack(bs).contiguous()


@pytest.mark.parametrize("w,v,n1", [(8, 64, 5000000), (4, 50, 1000003), (8, 300, 200000)])
def test_noalloc_matches_allocating_form(capi, w, v, n1):
    # Two calls in a row, the second on fewer rows: the scratch keeps the
    # first call's tile sums, which must not leak into the second.  5M rows
    # are 1221 scan tiles, more than one block of tile sums.
    for n, seed in (n1, 0), (9000, 1):
        x = _rows(n, v, w, seed)
        mask, bitval = _masks(x, 1)
        want = _offsets(capi, "ibp_condensed_offsets", x, mask[0], bitval[0])
        got = _offsets(capi, "ibp_condensed_offsets_noalloc", x, mask[0], bitval[0])
        torch.cuda.synchronize()
        assert torch.equal(got, want), (w, v, n)
        sizes = torch.diff(got, prepend=torch.zeros(1, dtype=torch.int64, device="cuda"))
        assert int(sizes.min()) > 0 and int(sizes.max()) <= v * w


def test_noalloc_index_and_clusters(capi):
    x = _rows(300000, 64, 8, 2)
    k = 3
    mask, bitval = _masks(x, k)
    g = torch.Generator().manual_seed(5)
    for n in (250000, 70000):
        index = torch.randint(0, x.shape[0], (n,), generator=g).cuda()
        clusters = torch.randint(0, k, (x.shape[0],), generator=g, dtype=torch.int32).cuda()
        for idx in (index, index.to(torch.int32)):
            want = _offsets(capi, "ibp_condensed_offsets", x, mask, bitval, idx, clusters, k)
            got = _offsets(capi, "ibp_condensed_offsets_noalloc", x, mask, bitval, idx, clusters, k)
            torch.cuda.synchronize()
            assert torch.equal(got, want), (n, idx.dtype)


def test_noalloc_limits(capi):
    x = _rows(64, 64, 8, 3)
    mask, bitval = _masks(x, 1)
    index = torch.zeros(16, dtype=torch.int64, device="cuda")
    out = torch.zeros(16, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    # n above the limit is rejected before any GPU work is enqueued.
    too_many = 268435456 + 1
    with pytest.raises(capi.IBPError, match="allocation-free limit"):
        capi.call("ibp_condensed_offsets_noalloc", 8, x.data_ptr(), x.shape[0], 64, mask[0].data_ptr(),
                  bitval[0].data_ptr(), index.data_ptr(), 8, too_many, out.data_ptr(), None, stream)
    host_out = torch.zeros(64, dtype=torch.int64).pin_memory()
    with pytest.raises(capi.IBPError, match="device memory"):
        capi.call("ibp_condensed_offsets_noalloc", 8, x.data_ptr(), x.shape[0], 64, mask[0].data_ptr(),
                  bitval[0].data_ptr(), None, 0, 64, host_out.data_ptr(), None, stream)
    # n = 0 is a no-op.
    capi.call("ibp_condensed_offsets_noalloc", 8, x.data_ptr(), x.shape[0], 64, mask[0].data_ptr(),
              bitval[0].data_ptr(), None, 0, 0, out.data_ptr(), None, stream)
