"""The reference's own test inputs, run through the drop-in modules the way the
reference's tests call them (positional arguments), bit-exact against the
oracle and the original tensors.

- tests/compress_test.py:30-46 (`import ibp`): float32 [1200 x 128] uniform,
  its float64 view, dense int32 [100 x 100] and its int64 view, semidense int32
  [6 x 100] and its int64 view; float tensors are passed as they are.
- tests/decompression_thput.py:15-26,79-91 (`import ibp_cuda as ibp`): 100,000
  zero rows of 256 B .. 16 KiB (int32), masks with a target fraction of random
  bit positions, bitval = row 0, comp_len = (compressed / original) * V as a
  0-dim tensor, impl 0.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import ibp
    import ibp_cuda
    return ibp, ibp_cuda


def _int_view(t):
    return t.view({4: torch.int32, 8: torch.int64}[t.element_size()])


def _np_int(t):
    return _int_view(t.cpu()).numpy()


def _compress_test_inputs():
    g = torch.Generator().manual_seed(0)
    sp32 = torch.rand([1200, 128], dtype=torch.float32, generator=g).pin_memory()
    sp64 = sp32.clone().pin_memory().view(dtype=torch.float64)
    dense32 = torch.randint(2147483647, [100, 100], dtype=torch.int32, generator=g).pin_memory()
    dense64 = dense32.clone().pin_memory().view(dtype=torch.int64)
    semi32 = torch.randint(2147483647 // 100, [6, 100], dtype=torch.int32, generator=g).pin_memory()
    semi64 = semi32.clone().pin_memory().view(dtype=torch.int64)
    return {"sp32": sp32, "sp64": sp64, "dense32": dense32, "dense64": dense64, "semi32": semi32,
            "semi64": semi64}


@pytest.fixture(scope="module")
def inputs(mods):
    return _compress_test_inputs()


@pytest.mark.parametrize("name", ["sp32", "sp64", "dense32", "dense64", "semi32", "semi64"])
def test_compress_test_do_the_process(mods, oracle, inputs, name):
    """do_the_process (tests/compress_test.py:4-24) on each of its inputs."""
    ibp, _ = mods
    orig = inputs[name]
    mask, bitval = ibp.preprocess(orig)
    assert mask.dtype == _int_view(orig[:1]).dtype and mask.is_cuda
    comp_tensor = orig.detach().clone().pin_memory()
    comp_sizes = ibp.get_compress_size(comp_tensor, mask, bitval)
    torch.cuda.synchronize()
    bitmask = ibp.compress_inplace(comp_tensor, mask, bitval)
    torch.cuda.synchronize()
    decomp = ibp.decompress_fetch(comp_tensor, mask, bitval, bitmask, torch.device("cuda"))
    assert decomp.dtype == orig.dtype and decomp.is_cuda
    assert torch.equal(_int_view(orig.to(torch.device("cuda"))), _int_view(decomp))
    # against the oracle on the integer view of the same bits
    x = _np_int(orig)
    om, ob, _, _ = oracle.preprocess(x, None)
    assert np.array_equal(_np_int(mask).view(om.dtype), om) and np.array_equal(_np_int(bitval).view(ob.dtype), ob)
    want_sizes, _ = oracle.get_compress_size(x, om.view(x.dtype), ob.view(x.dtype))
    assert np.array_equal(comp_sizes.numpy(), want_sizes)
    want_comp, want_bm = oracle.compress(x, om.view(x.dtype), ob.view(x.dtype))
    assert np.array_equal(bitmask.cpu().numpy(), want_bm)
    assert np.array_equal(_np_int(comp_tensor), want_comp)


def _thput_mask(v, rate, rng):
    """make_mask_and_bitval (tests/decompression_thput.py:15-26): the first
    rate * 32V + V + 1 positions of a random permutation of the 32V bits."""
    bits = 32 * v
    count = min(bits, int(rate * bits + v + 1))
    pos = rng.permutation(bits)[:count]
    m = np.zeros(v, dtype=np.uint32)
    np.bitwise_or.at(m, pos // 32, (np.uint32(1) << (pos % 32).astype(np.uint32)))
    return torch.from_numpy(m.view(np.int32))


@pytest.mark.parametrize("size", [256, 1024, 4 * 1024, 16 * 1024])
def test_decompression_thput_round_trip(mods, oracle, size):
    """compress_decompress (tests/decompression_thput.py:39-79) at its sizes,
    100,000 rows, for four of its target fractions."""
    _, ibp = mods
    v = size // 4
    tensor = torch.zeros([100000, v], dtype=torch.int32).pin_memory()
    rng = np.random.default_rng(size)
    for rate in (0.125, 0.5, 0.9, 0.97):
        print("decompression_thput", size, rate, flush=True)
        mask = _thput_mask(v, rate, rng).to(torch.device("cuda"))
        bitval = tensor[0].clone().to(torch.device("cuda"))
        copy_tensor = tensor.detach().clone().pin_memory()
        comp_sizes = ibp.get_compress_size(copy_tensor, mask, bitval, None, None)
        torch.cuda.synchronize()
        og_size = copy_tensor.numel() * copy_tensor.element_size()
        comp_size = torch.sum(comp_sizes)
        bitmask = ibp.compress_inplace(copy_tensor, mask, bitval, None)
        torch.cuda.synchronize()
        decomp = ibp.decompress_fetch(copy_tensor, mask, bitval, bitmask, torch.device("cuda"),
                                      (comp_size / og_size) * tensor.shape[1], None, None, None, 0)
        torch.cuda.synchronize()
        assert torch.equal(tensor, decomp.to(torch.device("cpu")))
        # every row is the same zero row: its size and slot against the oracle
        m, b = mask.cpu().numpy(), bitval.cpu().numpy()
        row = tensor[:1].numpy()
        want_size = oracle.row_size(row[0], m, b)
        assert int(comp_sizes.min()) == int(comp_sizes.max()) == want_size
        want_slot, want_bm = oracle.compress(row, m, b)
        got = copy_tensor[:2].numpy()
        assert np.array_equal(got[0], want_slot[0]) and np.array_equal(got[1], want_slot[0])
        assert (int(bitmask[0]) & 1) == (int(want_bm[0]) & 1)
