"""The HIP path on the benchmark workloads at their stated sizes, bit-exact:

- GNN minibatch table, fp32 [2,449,029 x 100] as int64 [2,449,029 x 50]:
  masks, every row's compressed size, every slot and the bitmask against the
  oracle, in HBM and in pinned memory; whole 200,000-row minibatches fetched
  from both against the original rows;
- DLRM, 26 fp32 tables of dim 128 (rows log-uniform in [1e3, 1e7], 59.8M rows
  in all) as int64 [rows x 64]: one plan, one launch, eager, autotuned and
  replayed from a HIP graph with refilled indices, against x[idx] per table;
- LLM weight blocks, fp16 [4096 x 4096] and [4096 x 11008] as int64
  [4096 x 1024] / [4096 x 2752]: preprocess_kmeans(16) against the oracle,
  then per-cluster and mask-row-0 pack and fetch;
- serving: Poisson arrivals of requests of 1e3..2.5e5 rows over the GNN
  feature, GNN minibatch and DLRM sets, every result checked, and one request
  identical alone and in the mix.

The reference's semantics: src/decompress.cu:57-127 (gather), the round trip
of tests/decompression_thput.py:69-72, src/compress.cu:176-236 (pack),
include/preproc/ibp_preproc_host.cuh:24-115 and :119-284 (mask selection).
"""
import json
import math
import os

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


@pytest.fixture(scope="module")
def oracle_mt(oracle):
    oracle.set_threads(min(16, len(os.sched_getaffinity(0))))
    return oracle


def _u(a):
    return a.view(np.dtype("u%d" % a.itemsize))


def _packed_bits(bitmask, n):
    return np.unpackbits(bitmask.view(np.uint8), bitorder="little")[:n].astype(bool)


# ------------------------------------------------------------------ GNN minibatch table

@pytest.fixture(scope="module")
def gnn(ibp, oracle_mt):
    from invariantbitpacking_amd import workloads
    x = workloads.gnn_minibatch(device="cuda")
    xn = x.cpu().numpy()
    mask, bitval = ibp.preprocess(x, None)
    om, ob, _, _ = oracle_mt.preprocess(xn, None)
    torch.cuda.synchronize()
    return dict(x=x, xn=xn, mask=mask, bitval=bitval, om=om, ob=ob)


def test_gnn_minibatch_masks_sizes_slots(ibp, oracle_mt, gnn):
    x, xn = gnn["x"], gnn["xn"]
    n, v = xn.shape
    print("gnn minibatch table", xn.shape, flush=True)
    assert np.array_equal(_u(gnn["mask"].cpu().numpy()), gnn["om"])
    assert np.array_equal(_u(gnn["bitval"].cpu().numpy()), gnn["ob"])
    m, b = gnn["om"].view(np.int64), gnn["ob"].view(np.int64)
    sizes = ibp.get_compress_size(x, gnn["mask"], gnn["bitval"]).cpu().numpy()
    want_sizes, _ = oracle_mt.get_compress_size(xn, m, b)
    assert np.array_equal(sizes, want_sizes)
    assert np.all(sizes <= v * 8) and np.all(sizes % 8 == 0) and np.all(sizes > 0)
    want_comp, want_bm = oracle_mt.compress_mt(xn, m, b, min(16, len(os.sched_getaffinity(0))))
    assert np.array_equal(_packed_bits(want_bm, n), sizes < v * 8)
    # in HBM
    comp = x.clone()
    bm = ibp.compress_inplace(comp, gnn["mask"], gnn["bitval"])
    assert np.array_equal(bm.cpu().numpy(), want_bm)
    assert np.array_equal(_u(comp.cpu().numpy()), _u(want_comp))
    del comp
    # in pinned host memory
    pinned = torch.from_numpy(xn).pin_memory()
    bm = ibp.compress_inplace(pinned, gnn["mask"], gnn["bitval"])
    assert np.array_equal(bm.cpu().numpy(), want_bm)
    assert np.array_equal(_u(pinned.numpy()), _u(want_comp))
    gnn["comp_pinned"], gnn["bitmask"] = pinned, bm
    gnn["comp_len"] = int(sizes.sum() / (n * v * 8) * v)


def test_gnn_minibatch_fetch_whole_batches(ibp, oracle_mt, gnn):
    from invariantbitpacking_amd import workloads
    if "comp_pinned" not in gnn:
        pytest.skip("needs the compressed table of the previous test")
    x, pinned, bm = gnn["x"], gnn["comp_pinned"], gnn["bitmask"]
    n, v = x.shape
    dev = pinned.cuda()
    out = torch.empty((200000, v), dtype=torch.int64, device="cuda")
    for step in (0, 1, 63):
        idx = workloads.gnn_minibatch_indices(n, 200000, step).cuda()
        want = x[idx]
        for where, comp in (("hbm", dev), ("pinned", pinned)):
            print("minibatch", step, where, flush=True)
            out.fill_(-1)
            got = ibp.decompress_fetch(comp, gnn["mask"], gnn["bitval"], bm, torch.device("cuda"), gnn["comp_len"],
                                       idx, None, None, None, out)
            assert got is out
            assert torch.equal(out, want), (step, where)
    # the oracle's unpack of the GPU-written slots, on a row sample
    sample = np.random.default_rng(7).integers(0, n, size=5000)
    rows = oracle_mt.decompress_fetch(pinned.numpy(), gnn["om"].view(np.int64), gnn["ob"].view(np.int64),
                                      bm.cpu().numpy(), sample)
    assert np.array_equal(rows, gnn["xn"][sample])


# ------------------------------------------------------------------ DLRM, 26 tables

@pytest.fixture(scope="module")
def dlrm(ibp):
    from invariantbitpacking_amd import workloads
    tables = workloads.dlrm_tables(device="cuda")
    sets = []
    for t, x in enumerate(tables):
        mask, bitval = ibp.preprocess(x, None)
        comp = x.clone()
        bm = ibp.compress_inplace(comp, mask, bitval)
        L = ibp.get_single_compress_len(x, mask, bitval)
        sets.append(dict(x=x, comp=comp, mask=mask, bitval=bitval, bitmask=bm, comp_len=L))
    torch.cuda.synchronize()
    return sets


def test_dlrm_masks_small_tables_match_oracle(ibp, oracle_mt, dlrm):
    """Tables of up to 600K rows (15 of the 26): masks, sizes and slots
    against the oracle (the larger tables are covered by the round trip)."""
    checked = 0
    for t, s in enumerate(dlrm):
        if s["x"].shape[0] > 600000:
            continue
        xn = s["x"].cpu().numpy()
        om, ob, _, _ = oracle_mt.preprocess(xn, None)
        assert np.array_equal(_u(s["mask"].cpu().numpy()), om), t
        assert np.array_equal(_u(s["bitval"].cpu().numpy()), ob), t
        want_comp, want_bm = oracle_mt.compress_mt(xn, om.view(np.int64), ob.view(np.int64), 8)
        assert np.array_equal(s["bitmask"].cpu().numpy(), want_bm), t
        assert np.array_equal(_u(s["comp"].cpu().numpy()), _u(want_comp)), t
        checked += 1
    assert checked == sum(s["x"].shape[0] <= 600000 for s in dlrm) == 15


def test_dlrm_one_launch_graph_replays(ibp, dlrm):
    from invariantbitpacking_amd import workloads
    rows = [s["x"].shape[0] for s in dlrm]
    batches = [workloads.dlrm_indices(rows, 65536, seed) for seed in (101, 7001, 9001)]
    idx = [i.cuda() for i in batches[0]]
    outs = [torch.full((65536, 64), -1, dtype=torch.int64, device="cuda") for _ in dlrm]
    bf = ibp.BatchedFetch([dict(comp=s["comp"], mask=s["mask"], bitval=s["bitval"], bitmask=s["bitmask"],
                                index=i, out=o, comp_len=s["comp_len"]) for s, i, o in zip(dlrm, idx, outs)])
    desc = json.loads(bf.describe())
    print("dlrm plan:", desc, flush=True)
    assert desc["tasks"] == 26 and desc["segmented_blocks"] > 0  # one segmented launch
    bf.launch()
    torch.cuda.synchronize()
    for t, (s, i, o) in enumerate(zip(dlrm, idx, outs)):
        assert torch.equal(o, s["x"][i]), ("eager", t)
    ms = bf.autotune()
    assert ms > 0
    bf.capture_graph()
    for b, batch in enumerate(batches):
        for i, src in zip(idx, batch):
            i.copy_(src)
        for o in outs:
            o.fill_(-1)
        bf.graph_launch()
        torch.cuda.synchronize()
        for t, (s, i, o) in enumerate(zip(dlrm, idx, outs)):
            assert torch.equal(o, s["x"][i]), ("replay", b, t)
    bf.close()


# ------------------------------------------------------------------ LLM blocks, k-means

@pytest.fixture(scope="module")
def llm_blocks():
    from invariantbitpacking_amd import workloads
    layer = workloads.llm_weight_blocks(layers=1)[0]
    return [layer[0], layer[4]]  # [4096 x 1024] and [4096 x 2752] int64


@pytest.mark.parametrize("which", [0, 1])
def test_llm_block_kmeans16_matches_oracle(ibp, oracle_mt, llm_blocks, which):
    x = llm_blocks[which]
    xn = x.numpy()
    print("kmeans(16) on", tuple(x.shape), flush=True)
    om, ob, ocl = oracle_mt.kmeans(xn, 16, seed=1)
    pinned = x.clone().pin_memory()
    m, b, cl = ibp.preprocess_kmeans(pinned, 16, 0.85, seed=1)
    assert np.array_equal(cl.numpy(), ocl)
    assert np.array_equal(_u(m.cpu().numpy()), om) and np.array_equal(_u(b.cpu().numpy()), ob)
    # per-cluster pack and fetch
    want_comp, want_bm = oracle_mt.compress(xn, om.view(np.int64), ob.view(np.int64), clusters=ocl)
    comp = x.clone().pin_memory()
    bm = ibp.compress_inplace(comp, m, b, clusters=cl)
    assert np.array_equal(bm.cpu().numpy(), want_bm)
    assert np.array_equal(_u(comp.numpy()), _u(want_comp))
    out = ibp.decompress_fetch(comp, m, b, bm, torch.device("cuda"), clusters=cl.cuda())
    assert torch.equal(out.cpu(), x)
    # the reference's path: a [K, V] mask is applied through its row 0
    want0, want_bm0 = oracle_mt.compress(xn, om[0].view(np.int64), ob[0].view(np.int64))
    comp0 = x.clone().pin_memory()
    bm0 = ibp.compress_inplace(comp0, m, b)
    assert np.array_equal(bm0.cpu().numpy(), want_bm0) and np.array_equal(_u(comp0.numpy()), _u(want0))
    out0 = ibp.decompress_fetch(comp0, m, b, bm0, torch.device("cuda"))
    assert torch.equal(out0.cpu(), x)


# ------------------------------------------------------------------ serving mix

def test_serving_poisson_mix_every_result(ibp, gnn, dlrm):
    """Requests over the GNN feature set (pinned), the GNN minibatch table
    (HBM) and the 26 DLRM tables (HBM), n ~ logU[1e3, 2.5e5], Poisson
    arrivals at 1e3/s and 1e4/s: every result equals x[idx]; one request gives
    the same rows alone and in the mix."""
    import time
    from invariantbitpacking_amd import workloads
    feats, _ = workloads.gnn_features()
    fx = feats.cuda()
    fm, fb = ibp.preprocess(fx, 0.9)
    fcomp = feats.clone().pin_memory()
    fbm = ibp.compress_inplace(fcomp, fm, fb)
    gx = gnn["x"]
    gcomp = gx.clone()
    gbm = ibp.compress_inplace(gcomp, gnn["mask"], gnn["bitval"])
    srv = ibp.Server(hbm_budget_bytes=4 << 30, num_streams=4, max_batch_requests=32, max_batch_rows=1 << 20)
    sets = [(srv.register(fcomp, fm, fb, fbm, ibp.get_single_compress_len(fx, fm, fb)), fx),
            (srv.register(gcomp, gnn["mask"], gnn["bitval"], gbm), gx)]
    for s in dlrm:
        sets.append((srv.register(s["comp"], s["mask"], s["bitval"], s["bitmask"], s["comp_len"]), s["x"]))
    rng = np.random.default_rng(300)
    probe_set = 1
    probe_idx = rng.integers(0, gx.shape[0], size=30000)
    r = srv.submit(sets[probe_set][0], probe_idx)
    alone, _ = srv.wait(r)
    alone = alone.clone()
    srv.release(r)
    assert torch.equal(alone, gx[torch.from_numpy(probe_idx).cuda()])
    total = 0
    for rate, count in ((1e3, 60), (1e4, 200)):
        gaps = rng.exponential(1.0 / rate, size=count)
        plan = []
        for k in range(count):
            sk = int(rng.integers(0, len(sets)))
            n = int(math.exp(rng.uniform(math.log(1e3), math.log(2.5e5))))
            plan.append((sk, rng.integers(0, sets[sk][1].shape[0], size=n)))
        t0 = srv.clock_us()
        arrivals = t0 + 1e6 * np.cumsum(gaps)
        reqs = []
        for k, (sk, idx) in enumerate(plan):
            while srv.clock_us() < arrivals[k]:
                time.sleep(0)
            reqs.append(srv.submit(sets[sk][0], idx, arrival_us=float(arrivals[k])))
            if k == count // 2:
                mixed = srv.submit(sets[probe_set][0], probe_idx)
        print("rate %g: %d requests submitted" % (rate, count), flush=True)
        for (sk, idx), rq in zip(plan, reqs):
            got, lat = srv.wait(rq)
            assert lat > 0
            assert torch.equal(got, sets[sk][1][torch.from_numpy(idx).cuda()]), (rate, sk, len(idx))
            srv.release(rq)
            total += 1
        got, _ = srv.wait(mixed)
        assert torch.equal(got, alone), rate
        srv.release(mixed)
    st = srv.stats()
    assert st["completed"] == total + 3 and st["p99_us"] >= st["p50_us"] > 0
    srv.close()
