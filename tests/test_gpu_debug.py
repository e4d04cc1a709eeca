"""print_debug(True) reports what the reference's debug build prints
(include/preproc/ibp_preproc_host.cuh:66,89-98 for the threshold sweep,
include/decompress/ibp_decompress_host.cuh:52-61 for the fetch launch,
ibp_preproc_host.cuh:221-232 for the k-means centroids), with the values the
oracle computes."""
import re

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


def test_debug_sweep_report_matches_oracle(ibp, oracle, capfd):
    rng = np.random.default_rng(3)
    f = rng.normal(0, 0.1, size=(5000, 128)).astype(np.float32)
    f[rng.random(f.shape) < 0.3] = 0.0
    x = f.view(np.int64)
    om, ob, step, chosen_t = oracle.preprocess(x, None)
    ibp.print_debug(True)
    try:
        capfd.readouterr()
        ibp.preprocess(torch.from_numpy(x).cuda(), None)
        torch.cuda.synchronize()
        out = capfd.readouterr().out
    finally:
        ibp.print_debug(False)
    lines = re.findall(r"Threshold ([0-9.]+): Saved bits per element: ([0-9]+)", out)
    assert len(lines) == 7, out
    saved = [int(s) for _, s in lines]
    # the sweep's exact float32 values: 0.7f, then t = float(double(t) + 0.05)
    sweep, t = [], np.float32(0.7)
    while t <= np.float32(1.0):
        sweep.append(t)
        t = np.float32(float(t) + 0.05)
    cnt = oracle.count_bits(x)
    for (shown, s), t in zip(lines, sweep):
        assert float(shown) == pytest.approx(float(t), abs=0.006)
        m, b = oracle.create_mask(cnt, x.shape[0], x.shape[1], 8, t)
        assert oracle.check_feats(x, m, b) == int(s), (t, s)
    sel = re.findall(r"Selected threshold ([0-9.]+); compress_len ([0-9]+)", out)
    assert sel, out
    assert float(sel[-1][0]) == pytest.approx(chosen_t, abs=0.006)
    total = x.size * 64
    best = max(saved)
    assert int(sel[-1][1]) == int(x.shape[1] * (total - best) / total) + 1


def test_debug_fetch_and_kmeans_reports(ibp, capfd):
    x = torch.randint(0, 1 << 20, (4096, 64), dtype=torch.int64).cuda()
    mask, bitval = ibp.preprocess(x, 0.9)
    comp = x.clone()
    bm = ibp.compress_inplace(comp, mask, bitval)
    ibp.print_debug(True)
    try:
        capfd.readouterr()
        ibp.decompress_fetch(comp, mask, bitval, bm, torch.device("cuda"), None,
                             torch.arange(100, device="cuda"))
        ibp.preprocess_kmeans(x[:512, :16].contiguous(), 3, seed=1)
        torch.cuda.synchronize()
        out = capfd.readouterr().out
    finally:
        ibp.print_debug(False)
    assert re.search(r"decompress_fetch: wave per row group, impl -1, device rows, V=64, w=8, n=100", out), out
    cents = re.findall(r"Centroid ([0-9]+): Set bits ([0-9]+) of 1024 \([0-9.]+%\) for ([0-9]+) nodes", out)
    assert cents and sum(int(n) for _, _, n in cents) == 512, out
    assert "Finished clustering" in out
