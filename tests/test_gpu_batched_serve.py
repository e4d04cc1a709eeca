"""GPU parity of the batched (one launch per plan, HIP graph) and served
(continuous batching) forms: every result equals the plain decompress_fetch
of the same rows, which the parity suite pins to the oracle."""
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


def make_set(ibp, n, v, w, seed, where="device", threshold=None):
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 0.05, size=(n, v * w // 4)).astype(np.float32)
    f[rng.random(f.shape) < 0.3] = 0.0
    x = torch.from_numpy(f.view(np.dtype("i%d" % w)).copy())
    xd = x.cuda()
    mask, bitval = ibp.preprocess(xd, threshold)
    comp = xd.clone()
    bm = ibp.compress_inplace(comp, mask, bitval)
    L = ibp.get_single_compress_len(xd, mask, bitval)
    if where == "pinned":
        comp = comp.cpu().pin_memory()
    torch.cuda.synchronize()
    return dict(x=x, comp=comp, mask=mask, bitval=bitval, bitmask=bm, comp_len=L)


SETS = [(5000, 50, 8, "device"), (3000, 64, 8, "pinned"), (700, 1024, 8, "device"), (2000, 3, 4, "device"),
        (900, 130, 4, "pinned"), (4000, 64, 4, "device")]


@pytest.fixture(scope="module")
def sets(ibp):
    return [make_set(ibp, n, v, w, seed=i, where=where) for i, (n, v, w, where) in enumerate(SETS)]


def _tasks(sets, seed, itypes=(torch.int64, torch.int32)):
    g = torch.Generator().manual_seed(seed)
    tasks = []
    for k, s in enumerate(sets):
        n, v = s["x"].shape
        if k == 2:
            idx = None  # whole table, no index
            rows = n
        else:
            rows = int(torch.randint(1, 3 * n, (1,), generator=g))
            idx = torch.randint(0, n, (rows,), generator=g).to(itypes[k % len(itypes)]).cuda()
        out = torch.full((rows, v), -7, dtype=s["x"].dtype, device="cuda")
        tasks.append(dict(comp=s["comp"], mask=s["mask"], bitval=s["bitval"], bitmask=s["bitmask"], index=idx,
                          out=out, comp_len=s["comp_len"]))
    return tasks


def _check(sets, tasks):
    for s, t in zip(sets, tasks):
        want = s["x"] if t["index"] is None else s["x"][t["index"].long().cpu()]
        assert torch.equal(t["out"].cpu(), want)


def test_autotune_cache_and_config(ibp, sets, tmp_path):
    """autotune(cache=...) stores the best configuration under a key of the
    GPU and the task shapes; a plan of the same shapes takes it from the file
    without timing, and launches it bit-exactly; set_config rejects values out
    of range."""
    import json
    cache = str(tmp_path / "tune.json")
    tasks = _tasks(sets, 21)
    bf = ibp.BatchedFetch(tasks)
    ms = bf.autotune(cache=cache)
    with open(cache) as f:
        entries = json.load(f)
    assert len(entries) == 1
    (entry,) = entries.values()
    assert entry["config"] == bf.config() and entry["ms"] == round(ms, 5) and entry["tasks"] == len(sets)
    bf.launch()
    torch.cuda.synchronize()
    _check(sets, tasks)
    # same shapes, new buffers: a cache hit applies the stored configuration
    tasks2 = _tasks(sets, 21)
    bf2 = ibp.BatchedFetch(tasks2)
    default = bf2.config()
    assert bf2.autotune(cache=cache) == entry["ms"] and bf2.config() == entry["config"]
    bf2.launch()
    torch.cuda.synchronize()
    _check(sets, tasks2)
    # other shapes: a second entry
    bf3 = ibp.BatchedFetch(_tasks(sets[:2], 22))
    bf3.autotune(cache=cache)
    with open(cache) as f:
        assert len(json.load(f)) == 2
    # configurations by hand; a captured graph is dropped and captured again
    bf2.set_config(default)
    assert bf2.config() == default
    bf2.capture_graph()
    bf2.graph_launch()
    torch.cuda.synchronize()
    _check(sets, tasks2)
    for bad in ([0] + default[1:], default[:2] + [3] + default[3:], default[:5] + [2, default[6]]):
        with pytest.raises(RuntimeError):
            bf2.set_config(bad)
    for b in (bf, bf2, bf3):
        b.close()


def test_batched_matches_plain(ibp, sets):
    tasks = _tasks(sets, 1)
    bf = ibp.BatchedFetch(tasks)
    bf.launch()
    torch.cuda.synchronize()
    _check(sets, tasks)
    for t in tasks:
        t["out"].fill_(-1)
    ms = bf.autotune()
    assert ms > 0
    bf.launch()
    torch.cuda.synchronize()
    _check(sets, tasks)
    bf.close()


def test_batched_graph_replays(ibp, sets):
    """Graph replays with eager work and synchronisation between them, index
    buffers refilled in place; every replay is checked."""
    tasks = _tasks(sets, 2)
    bf = ibp.BatchedFetch(tasks)
    bf.capture_graph()
    g = torch.Generator().manual_seed(5)
    for rep in range(4):
        for s, t in zip(sets, tasks):
            if t["index"] is not None:
                n = s["x"].shape[0]
                t["index"].copy_(torch.randint(0, n, t["index"].shape, generator=g).to(t["index"].dtype))
            t["out"].fill_(rep)
        _ = torch.ones(1000, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        bf.graph_launch()
        torch.cuda.synchronize()
        _check(sets, tasks)
    assert '"graph": true' in bf.describe()
    bf.close()


@pytest.mark.parametrize("subset,itypes,segmented", [
    ((0, 1), (torch.int64,), True),                 # 8-byte rows <= 64 elements, int64 indices
    ((0, 1), (torch.int32,), True),                 # the same with int32 indices
    ((3, 5), (torch.int64,), True),                 # 4-byte rows <= 64 elements
    ((3, 5), (torch.int64, torch.int32), False),    # mixed index widths: persistent, one kind
    ((2, 4), (torch.int64,), False),                # wide rows only: persistent, one kind
])
def test_batched_launch_forms(ibp, sets, subset, itypes, segmented):
    """Each launch form of a plan (segmented per task, or persistent compiled
    for the kinds present) against the plain fetch, eagerly, after autotuning
    and over graph replays."""
    import json
    chosen = [sets[k] for k in subset]
    tasks = _tasks(chosen, 7, itypes)
    bf = ibp.BatchedFetch(tasks)
    assert (json.loads(bf.describe())["segmented_blocks"] > 0) == segmented
    bf.launch()
    torch.cuda.synchronize()
    _check(chosen, tasks)
    bf.autotune()
    bf.capture_graph()
    for rep in range(3):
        for t in tasks:
            t["out"].fill_(rep - 5)
        torch.cuda.synchronize()
        bf.graph_launch()
        torch.cuda.synchronize()
        _check(chosen, tasks)
    bf.close()


def test_batched_out_of_range_rows_are_zero(ibp, sets):
    s = sets[0]
    n, v = s["x"].shape
    idx = torch.tensor([0, n, 5, -3], dtype=torch.int64, device="cuda")
    out = torch.full((4, v), 9, dtype=torch.int64, device="cuda")
    bf = ibp.BatchedFetch([dict(comp=s["comp"], mask=s["mask"], bitval=s["bitval"], bitmask=s["bitmask"],
                                index=idx, out=out)])
    ibp.index_errors(reset=True)
    bf.launch()
    torch.cuda.synchronize()
    o = out.cpu()
    assert torch.equal(o[0], s["x"][0]) and torch.equal(o[2], s["x"][5])
    assert not o[1].any() and not o[3].any()
    assert ibp.index_errors(reset=True) == 2
    bf.close()


def test_server_results_equal_plain_alone_and_mixed(ibp, sets):
    srv = ibp.Server(hbm_budget_bytes=256 << 20, num_streams=3, max_batch_requests=16, max_batch_rows=1 << 18)
    ids = [srv.register(s["comp"], s["mask"], s["bitval"], s["bitmask"], s["comp_len"]) for s in sets]
    g = np.random.default_rng(0)
    # the same request alone...
    probe_idx = g.integers(0, sets[1]["x"].shape[0], size=777)
    r = srv.submit(ids[1], probe_idx)
    alone, lat = srv.wait(r)
    alone = alone.clone()
    srv.release(r)
    assert lat > 0
    # ...and in a mix of requests of every set submitted back to back
    reqs = []
    for k in range(60):
        sid = int(g.integers(0, len(sets)))
        n = sets[sid]["x"].shape[0]
        idx = g.integers(0, n, size=int(g.integers(1, 3000)))
        reqs.append((sid, idx, srv.submit(ids[sid], idx)))
        if k == 30:
            mixed_probe = srv.submit(ids[1], probe_idx)
    for sid, idx, rq in reqs:
        got, _ = srv.wait(rq)
        assert torch.equal(got.cpu(), sets[sid]["x"][torch.from_numpy(idx)])
        srv.release(rq)
    mixed, _ = srv.wait(mixed_probe)
    assert torch.equal(mixed, alone)
    assert torch.equal(alone.cpu(), sets[1]["x"][torch.from_numpy(probe_idx)])
    srv.release(mixed_probe)
    st = srv.stats()
    assert st["completed"] == 62 and st["p99_us"] >= st["p50_us"] > 0
    assert st["launches"] <= 62
    srv.close()


def test_server_admission_waits_for_free_hbm(ibp, sets):
    s = sets[0]
    v = s["x"].shape[1]
    srv = ibp.Server(hbm_budget_bytes=64 << 20, num_streams=2, max_batch_requests=8, max_batch_rows=1 << 16)
    sid = srv.register(s["comp"], s["mask"], s["bitval"], s["bitmask"], s["comp_len"])
    one = 4000 * v * 8
    srv.set_budget_override(one + 4096)  # room for one request's output at a time
    g = np.random.default_rng(1)
    idxs = [g.integers(0, s["x"].shape[0], size=4000) for _ in range(3)]
    reqs = [srv.submit(sid, i) for i in idxs]
    first, _ = srv.wait(reqs[0])
    with pytest.raises(RuntimeError):
        srv.wait(reqs[1], timeout_ms=200)  # cannot be admitted while the first output is held
    assert torch.equal(first.cpu(), s["x"][torch.from_numpy(idxs[0])])
    srv.release(reqs[0])
    for rq, i in zip(reqs[1:], idxs[1:]):
        got, _ = srv.wait(rq)
        assert torch.equal(got.cpu(), s["x"][torch.from_numpy(i)])
        srv.release(rq)
    st = srv.stats()
    assert st["waited_admissions"] >= 1 and st["peak_arena_bytes"] <= one + 4096
    wide = sets[2]  # 1024-element rows: 65536 of them (512 MiB) can never fit the 64 MiB arena
    wid = srv.register(wide["comp"], wide["mask"], wide["bitval"], wide["bitmask"], wide["comp_len"])
    with pytest.raises(RuntimeError, match="can never fit"):
        srv.submit(wid, np.zeros(1 << 16, dtype=np.int64))
    srv.close()


@pytest.mark.parametrize("w,shapes,itype", [
    (8, [(3000, 64, "device"), (2000, 50, "pinned"), (1000, 16, "device")], torch.int64),
    (4, [(1500, 48, "device"), (800, 64, "pinned")], torch.int32),
])
def test_batched_condensed_streams(ibp, w, shapes, itype):
    """Condensed streams of several tables in one plan (segmented launch)
    equal the original rows: eager, after autotuning and over graph replays,
    streams in HBM or pinned memory."""
    g = torch.Generator().manual_seed(11)
    chosen, tasks = [], []
    for k, (n, v, where) in enumerate(shapes):
        s = make_set(ibp, n, v, w, seed=40 + k)
        stream, offs = ibp.compress_condensed(s["x"].cuda(), s["mask"], s["bitval"])
        torch.cuda.synchronize()
        if where == "pinned":
            stream, offs = stream.cpu().pin_memory(), offs.cpu().pin_memory()
        rows = int(torch.randint(1, 2 * n, (1,), generator=g))
        idx = torch.randint(0, n, (rows,), generator=g).to(itype).cuda()
        out = torch.full((rows, v), -3, dtype=s["x"].dtype, device="cuda")
        chosen.append(s)
        tasks.append(dict(comp=stream, offsets=offs, vec_size=v, mask=s["mask"], bitval=s["bitval"], index=idx,
                          out=out, comp_len=s["comp_len"]))
    bf = ibp.BatchedFetch(tasks)
    assert '"segmented_blocks": 0' not in bf.describe()
    bf.launch()
    torch.cuda.synchronize()
    _check(chosen, tasks)
    bf.autotune()
    bf.capture_graph()
    for rep in range(3):
        for s, t in zip(chosen, tasks):
            t["index"].copy_(torch.randint(0, s["x"].shape[0], t["index"].shape, generator=g).to(itype))
            t["out"].fill_(rep)
        torch.cuda.synchronize()
        bf.graph_launch()
        torch.cuda.synchronize()
        _check(chosen, tasks)
    bf.close()


def test_batched_condensed_plan_rules(ibp, sets):
    """A plan holds only condensed tasks of rows <= 64 elements of one width."""
    s = sets[0]
    stream, offs = ibp.compress_condensed(s["x"].cuda(), s["mask"], s["bitval"])
    n, v = s["x"].shape
    out = torch.empty((n, v), dtype=s["x"].dtype, device="cuda")
    cond = dict(comp=stream, offsets=offs, vec_size=v, mask=s["mask"], bitval=s["bitval"], out=out)
    slot = dict(comp=s["comp"], mask=s["mask"], bitval=s["bitval"], bitmask=s["bitmask"], out=out.clone())
    with pytest.raises(RuntimeError, match="all condensed"):
        ibp.BatchedFetch([cond, slot])
    wide = sets[2]
    wstream, woffs = ibp.compress_condensed(wide["x"].cuda(), wide["mask"], wide["bitval"])
    wout = torch.empty(tuple(wide["x"].shape), dtype=wide["x"].dtype, device="cuda")
    with pytest.raises(RuntimeError, match="<= 64 elements"):
        ibp.BatchedFetch([dict(comp=wstream, offsets=woffs, vec_size=wide["x"].shape[1], mask=wide["mask"],
                               bitval=wide["bitval"], out=wout)])
