"""Benchmark of the IBP fetch path on one MI355X.

Workload (default): GNN minibatch gather-decompress -- fp32 node features
[2,449,029 x 100] (N(0, 0.5) clipped to [-4, 4], 25% zeros, 10% of columns
constant within 4096-row blocks) viewed as int64 [2,449,029 x 50], compressed
in place in HBM; one step = one decompress_fetch of 200,000 random rows into a
preallocated output.  Data are synthetic (seeded), generated on the GPU.

Prints ONE JSON line: the decompressed GB/s of the plain per-call fetch with
the compressed tensor resident in HBM (`value`), the fetch kernel's roofline
(algorithmic HBM bytes / kernel time, against 8 TB/s), a CPU baseline (the
oracle's single-threaded unpack of a bounded sample), and side figures: the
batched form (all minibatches in one persistent launch), the PCIe-resident
fetch against a measured pinned H2D copy ceiling, and call latency.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "decompressed GB/s into HBM (batched vs plain, bit-exact) + p99 fetch latency"
HBM_PEAK_GBPS = 8000.0



def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_region(fn, steps, warmup):
    """warmup calls, then `steps` calls bracketed by a synchronisation and HIP
    events on the current stream (the stream the kernels run on).  Returns the
    mean milliseconds per call from the events and the wall seconds of the
    region."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return s.elapsed_time(e) / steps, wall


def wait_until(t0, ta):
    """Sleeps (releasing the GIL) until 0.2 ms before t0 + ta seconds, then spins."""
    left = ta - (time.perf_counter() - t0)
    if left > 2e-4:
        time.sleep(left - 2e-4)
    while time.perf_counter() - t0 < ta:
        pass


def call_latencies(fn, calls):
    """Wall milliseconds of isolated calls (call issued -> results ready)."""
    out = []
    for _ in range(calls):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        out.append((time.perf_counter() - t0) * 1e3)
    return out


def h2d_ceiling_gbps(nbytes=256 << 20, reps=5):
    """Pinned host -> device copy of 256 MiB (the PCIe ceiling of this box)."""
    src = torch.empty(nbytes, dtype=torch.uint8).pin_memory()
    dst = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        dst.copy_(src, non_blocking=True)
    e.record()
    torch.cuda.synchronize()
    return nbytes * reps / (s.elapsed_time(e) / 1e3) / 1e9


def cpu_baseline(comp_np, om, ob, obm, idx_list, budget_s=10.0):
    """The oracle's single-threaded unpack of successive minibatches (at least
    one, until budget_s of CPU time)."""
    from oracle import oracle
    done_rows, t_total, batches = 0, 0.0, 0
    v, w = comp_np.shape[1], comp_np.itemsize
    for idx in idx_list:
        t0 = time.perf_counter()
        oracle.decompress_fetch(comp_np, om, ob, obm, idx)
        t_total += time.perf_counter() - t0
        done_rows += len(idx)
        batches += 1
        if t_total >= budget_s:
            break
    gbps = done_rows * v * w / t_total / 1e9
    return {"value": round(gbps, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": "%d minibatches x %d rows of the same compressed table, unpacked by the C oracle on one host "
                      "core (%.1f s)" % (batches, len(idx_list[0]), t_total)}


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"cpus": len(os.sched_getaffinity(0)), "model": model}


def cpu_compress_baseline(x_np, rows=100000):
    """Host-side compression by the C oracle on the first `rows` rows:
    preprocess (threshold sweep) on one core, pack on one core and on every
    core available to the process (OpenMP)."""
    from oracle import oracle
    sample = np.ascontiguousarray(x_np[:rows])
    nbytes = sample.nbytes
    oracle.set_threads(1)
    t0 = time.perf_counter()
    m, b, _, _ = oracle.preprocess(sample, None)
    t_pre = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.compress(sample, m, b)
    t_1 = time.perf_counter() - t0
    threads = len(os.sched_getaffinity(0))
    oracle.compress_mt(sample, m, b, threads)  # warm the thread pool
    t0 = time.perf_counter()
    oracle.compress_mt(sample, m, b, threads)
    t_mt = time.perf_counter() - t0
    return {"sample": "%d rows x %d int64 (%.1f MB)" % (sample.shape[0], sample.shape[1], nbytes / 1e6),
            "preprocess_1core_gbps": round(nbytes / t_pre / 1e9, 3),
            "compress_1core_gbps": round(nbytes / t_1 / 1e9, 3),
            "compress_all_cores_gbps": round(nbytes / t_mt / 1e9, 3), "threads": threads}


def algorithmic_hbm_bytes(sizes_sel, n, v, w):
    """Per launch: compressed bytes read + decompressed bytes written + int64
    index + bitmask words (ceil(n/32)*4)."""
    return int(sizes_sel) + n * v * w + n * 8 + (n + 31) // 32 * 4


def run_batched(ibp, comp, mask, bitval, bitmask, comp_len, idx, x, batch, v):
    """All minibatches of the list in ONE persistent launch (autotuned,
    replayed from a HIP graph)."""
    dec_bytes = batch * v * 8
    outs = [torch.empty((batch, v), dtype=torch.int64, device="cuda") for _ in idx]
    bf = ibp.BatchedFetch([dict(comp=comp, mask=mask, bitval=bitval, bitmask=bitmask, index=i, out=o,
                                comp_len=comp_len) for i, o in zip(idx, outs)])
    bf.autotune()
    bf.capture_graph()
    bms, _ = timed_region(bf.graph_launch, 10, 2)
    batched = len(idx) * dec_bytes / (bms / 1e3) / 1e9
    exact = all(bool(torch.equal(o, x[i])) for o, i in zip(outs, idx))
    desc = bf.describe()
    log("batched: %d minibatches/launch %.3f ms, %.1f GB/s (%s)" % (len(idx), bms, batched, desc))
    bf.close()
    return {"decompressed_gbps": round(batched, 2), "minibatches_per_launch": len(idx),
            "ms_per_launch": round(bms, 4), "bitexact": exact, "plan": json.loads(desc)}


def run_pcie(ibp, comp, mask, bitval, bitmask, comp_len, idx, x, outs, sel, traffic=None):
    """The GNN minibatch with the compressed table in pinned
    host memory, the kernel reading it over PCIe; every minibatch of the index
    list fetched once (timed), each output checked.  Roofline: compressed
    bytes per launch over kernel time against a measured pinned H2D copy;
    `traffic` is the host-read bytes per launch from counters (or null)."""
    comp_h = comp.cpu().pin_memory()
    ceiling = h2d_ceiling_gbps()
    step = [0]

    def fetch_h():
        k = step[0] % len(idx)
        ibp.decompress_fetch(comp_h, mask, bitval, bitmask, torch.device("cuda"), comp_len, idx[k], None, None, None,
                             outs[k])
        step[0] += 1
    hk, _ = timed_region(fetch_h, len(idx), 1)
    exact = sum(bool(torch.equal(o, x[i])) for o, i in zip(outs, idx))
    pcie_bytes = float(np.mean(sel))
    dec = outs[0].numel() * outs[0].element_size()
    achieved = pcie_bytes / (hk / 1e3) / 1e9
    res = {"decompressed_gbps": round(dec / (hk / 1e3) / 1e9, 2),
           "compressed_pcie_gbps": round(achieved, 2),
           "h2d_copy_ceiling_gbps": round(ceiling, 2),
           "frac_of_ceiling": round(achieved / ceiling, 3),
           "ms_per_call": round(hk, 4), "minibatches": len(idx), "bitexact_minibatches": exact,
           "bitexact": exact == len(idx),
           "roofline": {"bound": "pcie", "achieved": round(achieved, 2), "peak": round(ceiling, 2), "unit": "GB/s",
                        "frac": round(achieved / ceiling, 4),
                        "peak_what": "pinned 256 MiB hipMemcpy H2D, measured in this run",
                        "algorithmic_bytes_per_launch": int(pcie_bytes),
                        "traffic": None if not traffic else traffic.get("host_read_bytes"),
                        "traffic_what": None if not traffic else traffic.get("host_read_what")}}
    res["transfer_baselines"] = pcie_transfer_baselines(ibp, comp_h, mask, bitval, bitmask, comp_len, idx, x,
                                                        outs[0], max(3, min(16, len(idx))))
    del comp_h
    log("pcie: %s" % res)
    return res


def pcie_transfer_baselines(ibp, comp_h, mask, bitval, bitmask, comp_len, idx, x, out, steps):
    """The reference's own ways of moving the same rows over PCIe, on the same
    minibatches (decompressed GB/s = rows delivered to HBM per second):
      uva_raw_gather   uncompressed rows in pinned memory gathered by the GPU
                       (tests/decompression_thput2.py:40-49; impl 2 with an
                       all-zero bitmask);
      cpu_gather_copy  rows gathered by the host cores into a pinned buffer,
                       then one H2D copy (tests/copy_test.cu:113-121);
      fake_fetch       the compressed table read as the fetch reads it, no
                       unpack (impl 3, tests/dlrm_comp.py:95-97)."""
    dec = out.numel() * out.element_size()
    step = [0]
    raw_h = x.cpu().pin_memory()
    zero = torch.zeros_like(bitmask)

    def uva():
        ibp.decompress_fetch(raw_h, mask, bitval, zero, torch.device("cuda"), comp_len, idx[step[0] % len(idx)],
                             None, None, 2, out)
        step[0] += 1
    u_ms, _ = timed_region(uva, steps, 1)
    u_exact = bool(torch.equal(out, x[idx[(step[0] - 1) % len(idx)]]))
    idx_h = [i.cpu() for i in idx]
    stage = torch.empty(out.shape, dtype=out.dtype).pin_memory()

    def host_gather():
        i = idx_h[step[0] % len(idx_h)]
        torch.index_select(raw_h, 0, i, out=stage)
        out.copy_(stage, non_blocking=True)
        torch.cuda.synchronize()
        step[0] += 1
    host_gather()
    t0 = time.perf_counter()
    for _ in range(steps):
        host_gather()
    h_ms = (time.perf_counter() - t0) * 1e3 / steps
    h_exact = bool(torch.equal(out, x[idx[(step[0] - 1) % len(idx)]]))
    del raw_h, stage

    def fake():
        ibp.decompress_fetch(comp_h, mask, bitval, bitmask, torch.device("cuda"), comp_len, idx[step[0] % len(idx)],
                             None, None, 3, out)
        step[0] += 1
    f_ms, _ = timed_region(fake, steps, 1)
    return {"uva_raw_gather": {"ms_per_call": round(u_ms, 4), "decompressed_gbps": round(dec / (u_ms / 1e3) / 1e9, 2),
                               "bitexact": u_exact},
            "cpu_gather_copy": {"ms_per_call": round(h_ms, 4), "decompressed_gbps": round(dec / (h_ms / 1e3) / 1e9, 2),
                                "threads": torch.get_num_threads(), "bitexact": h_exact},
            "fake_fetch": {"ms_per_call": round(f_ms, 4), "decompressed_gbps": round(dec / (f_ms / 1e3) / 1e9, 2)}}


def _pmc_csv(outdir):
    """{kernel name: {counter: [value per dispatch]}} from the
    counter_collection CSV files rocprofv3 wrote under outdir."""
    import csv
    import glob
    per = {}
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                key = (row.get("Kernel_Name", ""), row.get("Counter_Name", ""), row.get("Dispatch_Id", ""))
                per[key] = per.get(key, 0.0) + float(row.get("Counter_Value", "nan"))
    out = {}
    for (k, c, _), val in per.items():
        out.setdefault(k, {}).setdefault(c, []).append(val)
    return out


def _pmc_pass(counters, ops, args, regex):
    """Runs `bench.py --workload kernels --ops ops` under `rocprofv3 --pmc
    counters` in a child process; returns {kernel: {counter: [values]}} or
    raises with the reason."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        raise RuntimeError("rocprofv3 not found")
    out = tempfile.mkdtemp(prefix="ibp_pmc_", dir="/tmp")
    cmd = ["timeout", "-s", "KILL", "150", prof, "--pmc"] + counters + [
        "--kernel-include-regex", regex, "--output-format", "csv", "-d", out, "-o", "pmc", "--",
        sys.executable, os.path.abspath(__file__), "--workload", "kernels", "--steps", "3", "--ops", ops,
        "--scale", str(args.scale), "--batch", str(args.batch)]
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("rocprofv3 exited %d: %s" % (r.returncode, r.stdout[-300:]))
    vals = _pmc_csv(out)
    shutil.rmtree(out, ignore_errors=True)
    if not vals:
        raise RuntimeError("no counter rows: %s" % r.stdout[-300:])
    return vals


def _mean_of(vals, kernel_part, counter):
    xs = [v for k, cs in vals.items() if kernel_part in k for c, vs in cs.items() if c == counter for v in vs]
    if not xs:
        raise RuntimeError("no %s values for a kernel named *%s*" % (counter, kernel_part))
    return float(np.mean(xs)), len(xs)


def measure_traffic(args):
    """DRAM and host-read bytes per launch of the headline fetch, measured live
    by rocprofv3 counter passes over child runs of this script (started before
    this process touches the GPU).  Each pass runs the plain fetch of one
    200,000-row minibatch beside a calibration kernel of known size: the
    raw-row kernel copying the whole uncompressed table in order (rows x 400
    bytes read and written) for HBM, and reading a 256 MiB pinned table in
    order for host memory.
      hbm_bytes: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the fetch,
        each scaled by known bytes / counter of the calibration copy;
      hbm_read_bytes_by_size: 32, 64 and 128 bytes times the fetch's
        TCC_EA0_RDREQ_{32B,64B,128B} (a cross-check without scaling);
      host_read_bytes: (TCC_EA0_RDREQ - TCC_EA0_RDREQ_DRAM) of the fetch from
        pinned memory -- L2 read requests that leave for host memory -- times
        the bytes per such request of the 256 MiB pinned read.
    Returns a dict, with the reason in "*_error" for a pass that failed."""
    from invariantbitpacking_amd import workloads
    res = {}
    t0 = time.perf_counter()
    rows = workloads._rows(2449029, args.scale)
    table = rows * 50 * 8
    kern = "fetch_small_kernel|fetch_baseline_kernel"
    sizes = ["TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum"]

    def by_size(vals, part):
        n32, _ = _mean_of(vals, part, sizes[1])
        n64, _ = _mean_of(vals, part, sizes[2])
        n128, _ = _mean_of(vals, part, sizes[3])
        return 32 * n32 + 64 * n64 + 128 * n128
    try:
        ops = "fetch_hbm,seq_copy_hbm"
        f = _pmc_pass(["FETCH_SIZE"], ops, args, kern)
        w_ = _pmc_pass(["WRITE_SIZE"], ops, args, kern)
        cf, _ = _mean_of(f, "fetch_baseline", "FETCH_SIZE")
        cw, _ = _mean_of(w_, "fetch_baseline", "WRITE_SIZE")
        ff, nf = _mean_of(f, "fetch_small_kernel", "FETCH_SIZE")
        fw, _ = _mean_of(w_, "fetch_small_kernel", "WRITE_SIZE")
        fs, ws = table / (cf * 1024), table / (cw * 1024)
        res["fetch_size_kib"], res["write_size_kib"] = round(ff, 1), round(fw, 1)
        res["fetch_scale"], res["write_scale"] = round(fs, 4), round(ws, 4)
        res["hbm_read_bytes"], res["hbm_write_bytes"] = int(ff * 1024 * fs), int(fw * 1024 * ws)
        res["hbm_bytes"] = res["hbm_read_bytes"] + res["hbm_write_bytes"]
        res["hbm_what"] = ("FETCH_SIZE x %.3f + WRITE_SIZE x %.3f KiB per fetch_small_kernel dispatch (%d dispatches); "
                           "scales from the raw-row kernel copying the %d-byte table in order" % (fs, ws, nf, table))
        z = _pmc_pass(sizes, ops, args, kern)
        res["hbm_read_bytes_by_size"] = int(by_size(z, "fetch_small_kernel"))
        res["calibration_read_bytes_by_size"] = int(by_size(z, "fetch_baseline"))
    except Exception as e:  # noqa: BLE001 -- reported, never fatal
        res["hbm_error"] = str(e)[:300]
    if not args.no_pcie:
        try:
            ops = "fetch_pinned,sequential_read_pinned_256MiB"
            c = ["TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_DRAM_sum"]
            p = _pmc_pass(c, ops, args, kern)
            sr, _ = _mean_of(p, "fetch_baseline", c[0])
            sd, _ = _mean_of(p, "fetch_baseline", c[1])
            fr, nf = _mean_of(p, "fetch_small_kernel", c[0])
            fd, _ = _mean_of(p, "fetch_small_kernel", c[1])
            per_req = (256 << 20) / max(1.0, sr - sd)
            res["host_read_bytes"] = int((fr - fd) * per_req)
            res["host_read_what"] = ("(TCC_EA0_RDREQ - TCC_EA0_RDREQ_DRAM) x %.1f bytes per request, per fetch "
                                     "dispatch from pinned memory (%d dispatches); bytes per request from a "
                                     "sequential read of a 256 MiB pinned table" % (per_req, nf))
            z = _pmc_pass(sizes, ops, args, kern)
            res["pinned_fetch_read_bytes_by_size"] = int(by_size(z, "fetch_small_kernel"))
        except Exception as e:  # noqa: BLE001
            res["host_read_error"] = str(e)[:300]
    res["seconds"] = round(time.perf_counter() - t0, 1)
    log("traffic: %s" % res)
    return res


def run_gnn_minibatch(args, traffic=None):
    """The headline: GNN minibatch table compressed in HBM, one plain
    decompress_fetch of 200,000 random rows per step, each of the 64
    minibatches into its own preallocated output (every output of the timed
    region is checked afterwards)."""
    import invariantbitpacking_amd as ibp
    from invariantbitpacking_amd import workloads

    t0 = time.perf_counter()
    x = workloads.gnn_minibatch(scale=args.scale, device="cuda")
    num_rows, v = x.shape
    w = 8
    log("data %s generated in %.1fs" % (tuple(x.shape), time.perf_counter() - t0))
    warm = x[:4096].clone()  # first calls load the kernels: keep that out of the timings
    wm, wb = ibp.preprocess(warm, None)
    ibp.compress_inplace(warm, wm, wb)
    reps = 5

    def wall_ms(fn, before=None):
        """Median wall milliseconds of `reps` synchronised calls of fn
        (`before`, untimed, runs ahead of each)."""
        ts = []
        for _ in range(reps):
            if before is not None:
                before()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))

    mask, bitval = ibp.preprocess(x, None)
    t_pre = wall_ms(lambda: ibp.preprocess(x, None))
    sizes = ibp.get_compress_size(x, mask, bitval)
    t_size = wall_ms(lambda: ibp.get_compress_size(x, mask, bitval))
    comp = torch.empty_like(x)
    t_pack = wall_ms(lambda: ibp.compress_inplace(comp, mask, bitval), before=lambda: comp.copy_(x))
    comp.copy_(x)
    bitmask = ibp.compress_inplace(comp, mask, bitval)
    tb = x.numel() * w
    gpu_compress = {"preprocess_gbps": round(tb / t_pre / 1e6, 1), "preprocess_ms": round(t_pre, 3),
                    "get_compress_size_gbps": round(tb / t_size / 1e6, 1), "get_compress_size_ms": round(t_size, 3),
                    "compress_inplace_gbps": round(tb / t_pack / 1e6, 1), "compress_inplace_ms": round(t_pack, 3),
                    "what": "whole %.1f MB table in HBM: median wall clock of %d synchronised calls each "
                            "(compress_inplace on a fresh copy each time)" % (tb / 1e6, reps)}
    comp_len = ibp.get_single_compress_len(x, mask, bitval)
    torch.cuda.synchronize()
    packed = int(torch.count_nonzero((sizes < v * w)))
    log("compressed: %d/%d rows packed, comp_len %d, ratio %.4f" %
        (packed, num_rows, comp_len, num_rows * v * w / float(sizes.sum())))

    batch = min(args.batch, num_rows)
    nsets = 64
    idx_cpu = [workloads.gnn_minibatch_indices(num_rows, batch, s) for s in range(nsets)]
    idx = [i.cuda() for i in idx_cpu]
    out = torch.empty((batch, v), dtype=torch.int64, device="cuda")
    step = [0]

    def fetch():
        ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), comp_len, idx[step[0] % nsets], None,
                             None, None, out)
        step[0] += 1

    # -- plain per-call fetch into one preallocated output, compressed table
    # resident in HBM (the metric)
    kern_ms, wall = timed_region(fetch, args.steps, args.warmup)
    dec_bytes = batch * v * w
    value = args.steps * dec_bytes / wall / 1e9
    # every minibatch of the timed sequence fetched again, untimed, and
    # compared after its own call
    timed = sorted({s % nsets for s in range(args.warmup, args.warmup + args.steps)})
    exact = 0
    for k in timed:
        ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), comp_len, idx[k], None, None, None,
                             out)
        exact += bool(torch.equal(out, x[idx[k]]))
    lat = call_latencies(fetch, max(200, args.steps))
    # side figure: a separate output per minibatch (5.1 GB of outputs), so the
    # rows written are not the ones the previous call left in the caches
    outs = [torch.empty((batch, v), dtype=torch.int64, device="cuda") for _ in range(nsets)]
    step_r = [0]

    def fetch_rot():
        k = step_r[0] % nsets
        ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), comp_len, idx[k], None, None, None,
                             outs[k])
        step_r[0] += 1
    rot_ms, rot_wall = timed_region(fetch_rot, args.steps, args.warmup)
    rot_exact = sum(bool(torch.equal(outs[k], x[idx[k]])) for k in
                    sorted({s % nsets for s in range(args.warmup, args.warmup + args.steps)}))
    sel = [int(sizes[i].sum()) for i in idx]
    traffic_alg = algorithmic_hbm_bytes(np.mean(sel), batch, v, w)
    achieved = traffic_alg / (kern_ms / 1e3) / 1e9
    log("plain: %.4f ms/call (events), %.1f GB/s decompressed, kernel-side %.1f GB/s HBM, %d/%d minibatches exact"
        % (kern_ms, value, achieved, exact, len(timed)))

    batched = None
    if not args.no_batched:
        batched = run_batched(ibp, comp, mask, bitval, bitmask, comp_len, idx, x, batch, v)
        batched["vs_plain"] = round(batched["decompressed_gbps"] / value, 3)
    pcie = None
    if not args.no_pcie:
        pcie = run_pcie(ibp, comp, mask, bitval, bitmask, comp_len, idx, x, outs, sel, traffic)
    del outs
    cpu = cpu_compress = None
    if not args.no_cpu:
        comp_np = comp.cpu().numpy()
        om = mask.cpu().numpy().view(np.uint64)
        ob = bitval.cpu().numpy().view(np.uint64)
        obm = bitmask.cpu().numpy()
        cpu = cpu_baseline(comp_np, om, ob, obm, [i.numpy() for i in idx_cpu], budget_s=args.cpu_seconds)
        log("cpu baseline: %s" % cpu)
        cpu_compress = cpu_compress_baseline(x[:100000].cpu().numpy())
        log("cpu compression: %s" % cpu_compress)

    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": "gnn_minibatch: %d x %d int64 (fp32 [%d x 100]) compressed in HBM, %d random rows "
                               "per step, decompress_fetch into preallocated outputs" % (num_rows, v, num_rows, batch),
                   "rows": num_rows, "vec_size": v, "batch": batch, "packed_rows": packed, "comp_len": comp_len},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": None if not traffic else traffic.get("hbm_bytes"),
                     "traffic_what": None if not traffic else traffic.get("hbm_what"),
                     "algorithmic_bytes_per_launch": traffic_alg, "kernel_ms": round(kern_ms, 5)},
        "cpu_baseline": cpu,
        "bitexact": exact == len(timed) and rot_exact == len(timed) and (batched is None or batched["bitexact"]) and
                    (pcie is None or pcie["bitexact"]),
        "bitexact_minibatches": "%d of %d minibatches of the timed sequence (each fetched again into the output "
                                "and compared)" % (exact, len(timed)),
        "separate_outputs": {"what": "the same steps, each minibatch into its own preallocated output",
                             "decompressed_gbps": round(args.steps * dec_bytes / rot_wall / 1e9, 2),
                             "kernel_ms": round(rot_ms, 5), "bitexact": rot_exact == len(timed)},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4),
                       "what": "isolated call: decompress_fetch issued -> rows ready (host wall clock), %d calls" % len(lat)},
        "batched": batched,
        "pcie_resident": pcie,
        "compression": {"gpu": gpu_compress, "cpu": cpu_compress},
        "host": host_info(),
    }
    if traffic:
        res["counters"] = {k: v for k, v in traffic.items() if k not in ("hbm_bytes", "host_read_bytes")}
    res["set"] = dict(x=x, comp=comp, mask=mask, bitval=bitval, bitmask=bitmask, comp_len=comp_len)
    return res


def _compress_set(ibp, x):
    mask, bitval = ibp.preprocess(x, None)
    sizes = ibp.get_compress_size(x, mask, bitval)
    comp = x.clone()
    bitmask = ibp.compress_inplace(comp, mask, bitval)
    comp_len = ibp.get_single_compress_len(x, mask, bitval)
    return dict(x=x, mask=mask, bitval=bitval, sizes=sizes, comp=comp, bitmask=bitmask, comp_len=comp_len)


def run_dlrm(args, condensed=True):
    """26 embedding tables (rows log-uniform in [1e3, 1e7], dim 128 fp32 as
    int64 [rows x 64]), each preprocessed separately; a batch gathers 65,536
    Zipf(1.05) rows per table.  Plain = 26 decompress_fetch calls; batched =
    one persistent launch (autotuned, HIP graph) for all 26 tables."""
    import invariantbitpacking_amd as ibp
    from invariantbitpacking_amd import workloads
    tables = workloads.dlrm_tables(scale=args.scale, device="cuda")
    sets = [_compress_set(ibp, x) for x in tables]
    torch.cuda.synchronize()
    rows = [x.shape[0] for x in tables]
    idx = [i.cuda() for i in workloads.dlrm_indices(rows, batch=args.batch)]
    outs = [torch.empty((args.batch, 64), dtype=torch.int64, device="cuda") for _ in tables]
    comp_bytes = sum(int(s["sizes"].sum()) for s in sets)
    raw_bytes = sum(r * 512 for r in rows)
    log("dlrm: %d tables, %d rows, ratio %.3f" % (len(tables), sum(rows), raw_bytes / comp_bytes))

    def plain():
        for s, i, o in zip(sets, idx, outs):
            ibp.decompress_fetch(s["comp"], s["mask"], s["bitval"], s["bitmask"], torch.device("cuda"),
                                 s["comp_len"], i, None, None, None, o)
    p_ms, p_wall = timed_region(plain, args.steps, args.warmup)
    exact = all(bool(torch.equal(o, s["x"][i])) for s, i, o in zip(sets, idx, outs))
    dec = len(tables) * args.batch * 512
    bf = ibp.BatchedFetch([dict(comp=s["comp"], mask=s["mask"], bitval=s["bitval"], bitmask=s["bitmask"], index=i,
                                out=o, comp_len=s["comp_len"]) for s, i, o in zip(sets, idx, outs)])
    bf.autotune()
    bf.capture_graph()
    for o in outs:
        o.fill_(0)
    b_ms, _ = timed_region(bf.graph_launch, args.steps, args.warmup)
    b_exact = all(bool(torch.equal(o, s["x"][i])) for s, i, o in zip(sets, idx, outs))
    sel = sum(int(s["sizes"][i].sum()) for s, i in zip(sets, idx))
    alg = sel + dec + len(tables) * (args.batch * 8 + (args.batch + 31) // 32 * 4)
    res = {"workload": "dlrm: 26 tables, %d rows total, 65536 Zipf(1.05) rows per table per batch" % sum(rows),
           "compression_ratio": round(raw_bytes / comp_bytes, 3),
           "plain_ms_per_batch": round(p_ms, 4), "plain_gbps": round(dec / (p_ms / 1e3) / 1e9, 1),
           "batched_ms_per_batch": round(b_ms, 4), "batched_gbps": round(dec / (b_ms / 1e3) / 1e9, 1),
           "batched_vs_plain": round(p_ms / b_ms, 3), "batched_hbm_alg_gbps": round(alg / (b_ms / 1e3) / 1e9, 1),
           "plan": json.loads(bf.describe()), "bitexact": exact and b_exact}
    bf.close()
    if condensed:
        res["condensed"] = run_dlrm_condensed(ibp, args, sets, idx, outs, raw_bytes)
    res["sets"] = sets
    return res


def batched_condensed(ibp, args, sets, idx, outs, streams, dec, steps):
    """The 26 condensed streams fetched by one graph-replayed launch of an
    autotuned plan."""
    bf = ibp.BatchedFetch([dict(comp=st, offsets=off, vec_size=64, mask=s["mask"], bitval=s["bitval"], index=i,
                                out=o, comp_len=s["comp_len"]) for (st, off), s, i, o in zip(streams, sets, idx, outs)])
    bf.autotune()
    bf.capture_graph()
    for o in outs:
        o.fill_(-1)
    b_ms, _ = timed_region(bf.graph_launch, steps, 1)
    exact = all(bool(torch.equal(o, s["x"][i])) for s, i, o in zip(sets, idx, outs))
    bf.close()
    return {"ms_per_batch": round(b_ms, 4), "decompressed_gbps": round(dec / (b_ms / 1e3) / 1e9, 1), "bitexact": exact}


def run_dlrm_condensed(ibp, args, sets, idx, outs, raw_bytes):
    """The same tables as condensed streams (compress_condensed): bytes kept,
    the fetch from a device-resident stream, and -- the deployment the stream
    exists for -- the fetch from pinned host memory over PCIe, next to the
    uncompressed pinned row gather (the reference's UVA baseline,
    tests/decompression_thput2.py:40-49)."""
    streams = []
    for s in sets:
        st, off = ibp.compress_condensed(s["x"], s["mask"], s["bitval"])
        streams.append((st, off))
    torch.cuda.synchronize()
    stream_bytes = sum(st.numel() * 8 for st, _ in streams)
    dec = len(sets) * args.batch * 512

    def fetch(src):
        for (st, off), s, i, o in zip(src, sets, idx, outs):
            ibp.decompress(st, off, s["mask"], s["bitval"], torch.device("cuda"), 64, i, o, s["comp_len"])
    d_ms, _ = timed_region(lambda: fetch(streams), args.steps, args.warmup)
    d_exact = all(bool(torch.equal(o, s["x"][i])) for s, i, o in zip(sets, idx, outs))
    res = {"stream_bytes": stream_bytes, "raw_bytes": raw_bytes, "host_footprint_ratio": round(raw_bytes / stream_bytes, 3),
           "device": {"ms_per_batch": round(d_ms, 4), "gbps": round(dec / (d_ms / 1e3) / 1e9, 1), "bitexact": d_exact}}
    res["device_batched"] = batched_condensed(ibp, args, sets, idx, outs, streams, dec, args.steps)
    log("dlrm condensed (device): %s" % res)
    if args.no_pcie:
        return res
    pinned = [(st.cpu().pin_memory(), off) for st, off in streams]
    del streams
    torch.cuda.synchronize()
    steps = max(3, args.steps // 8)
    p_ms, _ = timed_region(lambda: fetch(pinned), steps, 1)
    p_exact = all(bool(torch.equal(o, s["x"][i])) for s, i, o in zip(sets, idx, outs))
    lens = [off.diff(prepend=torch.zeros(1, dtype=torch.int64, device=off.device)) for _, off in pinned]
    pcie = sum(int(l[i].sum()) for l, i in zip(lens, idx))
    res["pinned"] = {"ms_per_batch": round(p_ms, 3), "decompressed_gbps": round(dec / (p_ms / 1e3) / 1e9, 2),
                     "compressed_pcie_gbps": round(pcie / (p_ms / 1e3) / 1e9, 2), "pcie_bytes_per_batch": pcie,
                     "bitexact": p_exact}
    res["pinned_batched"] = batched_condensed(ibp, args, sets, idx, outs, pinned, dec, steps)
    del pinned
    # uncompressed rows in pinned memory, gathered by the GPU (impl 2: raw rows, no unpack)
    raw = [s["x"].cpu().pin_memory() for s in sets]
    zero = [torch.zeros_like(s["bitmask"]) for s in sets]
    torch.cuda.synchronize()

    def gather():
        for r, z, s, i, o in zip(raw, zero, sets, idx, outs):
            ibp.decompress_fetch(r, s["mask"], s["bitval"], z, torch.device("cuda"), 64, i, None, None, 2, o)
    g_ms, _ = timed_region(gather, steps, 1)
    g_exact = all(bool(torch.equal(o, s["x"][i])) for s, i, o in zip(sets, idx, outs))
    res["raw_pinned_gather"] = {"ms_per_batch": round(g_ms, 3), "gbps": round(dec / (g_ms / 1e3) / 1e9, 2),
                                "bitexact": g_exact}
    res["pinned_vs_raw_gather"] = round(g_ms / p_ms, 3)
    log("dlrm condensed (pinned): %s" % {k: res[k] for k in ("pinned", "pinned_batched", "raw_pinned_gather", "pinned_vs_raw_gather")})
    del raw
    return res


def run_serve(args, groups=None):
    """Requests drawn from resident sets (GNN features, GNN minibatch table,
    DLRM tables), index counts log-uniform in [1e3, 2.5e5], Poisson arrivals
    at each rate.  Served: continuous batching over 4 streams.  Plain: one
    decompress_fetch per request in arrival order.  Both latencies run from
    the request's scheduled arrival to its completion; throughput is the
    decompressed bytes over first scheduled arrival -> last completion.  Every
    result of both forms is kept (the arena and the plain outputs hold a whole
    pass) and compared with x[idx] after the pass."""
    import invariantbitpacking_amd as ibp
    from invariantbitpacking_amd import workloads
    if groups is None:
        d1, _ = workloads.gnn_features()
        s1 = _compress_set(ibp, d1.cuda())
        s2 = _compress_set(ibp, workloads.gnn_minibatch(scale=args.scale, device="cuda"))
        s3 = [_compress_set(ibp, t) for t in workloads.dlrm_tables(scale=args.scale, device="cuda")]
        groups = [[s1], [s2], s3]
    # a request picks a workload uniformly, then one of its tensor sets
    sets = [st for g in groups for st in g]
    first = np.cumsum([0] + [len(g) for g in groups])
    torch.cuda.synchronize()
    rng = np.random.default_rng(300)
    passes = []
    for rate in list(args.rates) + [float("inf")]:
        nreq = 400 if rate == float("inf") else int(min(400, max(50, rate * 0.5)))
        arrivals = np.zeros(nreq) if rate == float("inf") else np.cumsum(rng.exponential(1.0 / rate, size=nreq))
        reqs = []
        for _ in range(nreq):
            grp = int(rng.integers(0, len(groups)))
            k = int(first[grp] + rng.integers(0, len(groups[grp])))
            n = int(np.exp(rng.uniform(np.log(1e3), np.log(2.5e5))))
            reqs.append((k, rng.integers(0, sets[k]["x"].shape[0], size=n)))
        passes.append((rate, arrivals, reqs))
    # The arena holds every output of a pass (they are checked after it), with
    # room for the 256-byte alignment of each region.
    need = max(sum(len(i) * sets[k]["x"].shape[1] * 8 + 256 for k, i in reqs) for _, _, reqs in passes)
    srv = ibp.Server(hbm_budget_bytes=int(need * 1.05) + (256 << 20), num_streams=4, max_batch_requests=64,
                     max_batch_rows=1 << 21)
    sids = [srv.register(s["comp"], s["mask"], s["bitval"], s["bitmask"], s["comp_len"]) for s in sets]
    # first launches load kernels and set up streams: keep that out of the latencies
    for sid, s in zip(sids, sets):
        w = np.arange(1000) % s["x"].shape[0]
        h = srv.submit(sid, w)
        srv.wait(h)
        srv.release(h)
        ibp.decompress_fetch(s["comp"], s["mask"], s["bitval"], s["bitmask"], torch.device("cuda"), s["comp_len"],
                             torch.from_numpy(w).cuda())
    torch.cuda.synchronize()
    results = []
    total_checked = 0
    all_exact = True
    for rate, arrivals, reqs in passes:
        nreq = len(reqs)
        dec = sum(len(i) * sets[k]["x"].shape[1] * 8 for k, i in reqs)
        # served: every request is submitted at its scheduled arrival (the
        # server stamps that time and its own completion time); the results
        # stay in the arena and are collected after the last submission.
        hs = []
        t0_us = srv.clock_us() + 1000.0  # the schedule starts 1 ms from now
        for (k, i), ta in zip(reqs, arrivals):
            due = t0_us + ta * 1e6
            left = (due - srv.clock_us()) / 1e6
            if left > 3e-4:
                time.sleep(left - 3e-4)
            while srv.clock_us() < due:
                pass
            hs.append(srv.submit(sids[k], i, arrival_us=due))
        done = []
        for h in hs:
            srv.wait(h, view=False)
            done.append(h)
        times = np.array([srv.request_times(h) for h in done])  # arrival, admitted, launch, issued, done (us)
        lat = (times[:, 4] - times[:, 0]) / 1e3
        served_wall = (times[:, 4].max() - (t0_us + arrivals[0] * 1e6)) / 1e6
        worst = times[int(np.argmax(lat))]
        ok = True
        for (k, i), h in zip(reqs, done):
            got, _ = srv.wait(h)
            ok = ok and bool(torch.equal(got, sets[k]["x"][torch.from_numpy(i).cuda()]))
            srv.release(h)
        # plain per-request calls, outputs kept for the same check
        plat, pouts = [], []
        t0 = time.perf_counter() + 1e-3
        for (k, i), ta in zip(reqs, arrivals):
            wait_until(t0, ta)
            s = sets[k]
            pouts.append(ibp.decompress_fetch(s["comp"], s["mask"], s["bitval"], s["bitmask"], torch.device("cuda"),
                                              s["comp_len"], torch.from_numpy(i).cuda()))
            torch.cuda.synchronize()
            plat.append((time.perf_counter() - t0 - ta) * 1e3)
        plain_wall = time.perf_counter() - t0 - arrivals[0]
        p_ok = all(bool(torch.equal(o, sets[k]["x"][torch.from_numpy(i).cuda()])) for (k, i), o in zip(reqs, pouts))
        del pouts
        total_checked += 2 * nreq
        all_exact = all_exact and ok and p_ok
        r = {"rate_per_s": "burst" if rate == float("inf") else rate, "requests": nreq,
             "served": {"p50_ms": round(float(np.percentile(lat, 50)), 3),
                        "p99_ms": round(float(np.percentile(lat, 99)), 3),
                        "gbps": round(dec / served_wall / 1e9, 2),
                        "slowest_ms": {"admit": round((worst[1] - worst[0]) / 1e3, 3),
                                       "to_launch": round((worst[2] - worst[1]) / 1e3, 3),
                                       "issue": round((worst[3] - worst[2]) / 1e3, 3),
                                       "to_done": round((worst[4] - worst[3]) / 1e3, 3)},
                        "bitexact": ok},
             "plain": {"p50_ms": round(float(np.percentile(plat, 50)), 3),
                       "p99_ms": round(float(np.percentile(plat, 99)), 3),
                       "gbps": round(dec / plain_wall / 1e9, 2), "bitexact": p_ok},
             "offered_gbps": None if rate == float("inf") else round(dec / (arrivals[-1] - arrivals[0]) / 1e9, 2)}
        r["served_vs_plain_gbps"] = round(r["served"]["gbps"] / r["plain"]["gbps"], 3)
        log("serve rate %g: %s" % (rate, r))
        results.append(r)
    st = srv.stats()
    srv.close()
    return {"workload": "serving: Poisson arrivals over %d resident sets of %d workloads, n ~ logU[1e3, 2.5e5]"
                        % (len(sets), len(groups)),
            "latency": "scheduled arrival -> completion, both forms", "rates": results,
            "results_checked": total_checked, "bitexact": all_exact,
            "server_stats": {k: (round(v, 2) if isinstance(v, float) else v) for k, v in st.items()}}


def run_pcie_rows(args):
    """What random rows from pinned memory can reach over PCIe, by row size:
    uncompressed int64 rows of 400, 512 and 1024 bytes (about 1 GB per table)
    gathered by decompress_fetch's raw-row path (impl 2, all-zero bitmask), 200,000
    uniform random rows per call, against a pinned 256 MiB H2D copy.  Reads
    from host memory go out in 128-byte lines: rows that are a whole number of
    aligned lines reach the copy rate, a 400-byte row at a 400-byte stride
    pays for 4-5 lines."""
    import invariantbitpacking_amd as ibp
    ceiling = h2d_ceiling_gbps()
    n = 200000
    g = torch.Generator().manual_seed(7)
    res = {"workload": "pcie_rows: 200000 random uncompressed rows from pinned memory per call",
           "h2d_copy_ceiling_gbps": round(ceiling, 2), "rows": []}
    for row_bytes in (400, 512, 1024):
        v = row_bytes // 8
        num = (1 << 30) // row_bytes
        table = torch.randint(-1 << 62, 1 << 62, (num, v), dtype=torch.int64, generator=g).pin_memory()
        zero = torch.zeros((num + 31) // 32, dtype=torch.int32, device="cuda")
        mask = torch.zeros(v, dtype=torch.int64, device="cuda")
        idx = [torch.randint(0, num, (n,), generator=g).cuda() for _ in range(4)]
        out = torch.empty((n, v), dtype=torch.int64, device="cuda")
        step = [0]

        def gather():
            ibp.decompress_fetch(table, mask, mask, zero, torch.device("cuda"), v, idx[step[0] % 4], None, None, 2,
                                 out)
            step[0] += 1
        ms, _ = timed_region(gather, 8, 2)
        exact = bool(torch.equal(out.cpu(), table[idx[(step[0] - 1) % 4].cpu()]))
        gbps = n * row_bytes / (ms / 1e3) / 1e9
        res["rows"].append({"row_bytes": row_bytes, "ms_per_call": round(ms, 4), "gbps": round(gbps, 2),
                            "frac_of_ceiling": round(gbps / ceiling, 3), "bitexact": exact})
        log("pcie rows %d B: %s" % (row_bytes, res["rows"][-1]))
        del table
    return res


def run_gnn_features(args, calls=64):
    """GNN features (65536 x 128 fp32 as int64 [65536 x 64], N(0, 0.1) with 30%
    zeros; preprocess at 0.9; compress_inplace in pinned host memory): calls
    of 4096 random rows, latency-bound.  For the table in HBM and in pinned
    memory: the isolated call's wall latency, `calls` back-to-back plain
    calls (HIP events), the same `calls` index sets fetched by ONE
    graph-replayed batched launch, and the host time of one call (the
    Python call and the C++ binding's checks and launch, no GPU wait).  Every output is checked."""
    import invariantbitpacking_amd as ibp
    from invariantbitpacking_amd import workloads
    x, _ = workloads.gnn_features()
    xd = x.cuda()
    mask, bitval = ibp.preprocess(xd, 0.9)
    pinned = x.clone().pin_memory()
    bm = ibp.compress_inplace(pinned, mask, bitval)
    cl = ibp.get_single_compress_len(xd, mask, bitval)
    comp = pinned.cuda()
    n, v = x.shape
    g = torch.Generator().manual_seed(1)
    idx = [torch.randint(0, n, (4096,), generator=g).cuda() for _ in range(calls)]
    outs = [torch.empty((4096, v), dtype=torch.int64, device="cuda") for _ in range(calls)]
    want = [xd[i] for i in idx]
    dev = torch.device("cuda")
    call_bytes = 4096 * v * 8
    res = {"workload": "gnn_features: 65536 x 64 int64 (fp32 [65536 x 128]), calls of 4096 random rows",
           "out_bytes_per_call": call_bytes, "calls": calls}
    for where, src in (("hbm", comp), ("pinned", pinned)):
        step = [0]

        def one():
            k = step[0] % calls
            ibp.decompress_fetch(src, mask, bitval, bm, dev, cl, idx[k], None, None, None, outs[k])
            step[0] += 1
        lat = call_latencies(one, max(1000, args.steps))
        p_ms, _ = timed_region(lambda: [one() for _ in range(calls)], max(3, args.steps // 4), 1)
        p_exact = all(bool(torch.equal(o, w)) for o, w in zip(outs, want))
        # host cost of one call: issue `calls` calls without waiting for the GPU
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            one()
        host_us = (time.perf_counter() - t0) / calls * 1e6
        torch.cuda.synchronize()
        for o in outs:
            o.fill_(-1)
        bf = ibp.BatchedFetch([dict(comp=src, mask=mask, bitval=bitval, bitmask=bm, index=i, out=o, comp_len=cl)
                               for i, o in zip(idx, outs)])
        bf.autotune()
        bf.capture_graph()
        for o in outs:
            o.fill_(-1)
        b_ms, _ = timed_region(bf.graph_launch, max(3, args.steps // 4), 1)
        b_exact = all(bool(torch.equal(o, w)) for o, w in zip(outs, want))
        bf.close()
        res[where] = {"p50_ms": round(float(np.percentile(lat, 50)), 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
                      "latency_calls": len(lat),
                      "plain_ms_per_call": round(p_ms / calls, 5),
                      "plain_gbps": round(calls * call_bytes / (p_ms / 1e3) / 1e9, 1),
                      "host_us_per_call": round(host_us, 2),
                      "batched_ms_per_%d_calls" % calls: round(b_ms, 4),
                      "batched_gbps": round(calls * call_bytes / (b_ms / 1e3) / 1e9, 1),
                      "batched_vs_plain": round(p_ms / b_ms, 3), "bitexact": p_exact and b_exact}
        log("gnn_features %s: %s" % (where, res[where]))
    res["set"] = dict(x=xd, comp=comp, mask=mask, bitval=bitval, bitmask=bm, comp_len=cl)
    return res


def run_kernels(args):
    """One short pass over every hot kernel on the GNN minibatch table, for
    rocprofv3 (each op `--steps` times, HIP events around each):
    preprocess (sweep), get_compress_size, compress_inplace, the plain fetch of
    200,000 random rows from HBM, the raw-row gather (impl 2) of the same rows
    from the UNCOMPRESSED table, torch's x[idx] gather, compress_condensed of
    the whole table and the fetch of the same rows from its stream, preprocess
    and k-means(16) of pinned host data (staged in HBM once), k-means(4) and
    k-means(16) of the whole table and k-means(16) of one 1024-wide and one
    2752-wide LLM block, a 256 MiB device copy
    (the counter calibration for HBM), the same fetch and raw gather from
    pinned memory, and a sequential read of a 256 MiB pinned table (every row
    in order, impl 2: the calibration for host-read counters)."""
    import invariantbitpacking_amd as ibp
    from invariantbitpacking_amd import workloads
    reps = max(1, args.steps)
    x = workloads.gnn_minibatch(scale=args.scale, device="cuda")
    n, v = x.shape
    dev = torch.device("cuda")
    res = {"workload": "kernels: gnn_minibatch %d x %d int64, %d reps per op" % (n, v, reps)}

    ops = set(args.ops.split(",")) if args.ops else None
    want = lambda name: ops is None or name in ops  # noqa: E731

    def ev(name, fn, nbytes=None):
        if ops is not None and name not in ops:
            return
        fn()  # first call: kernel load, allocations
        torch.cuda.synchronize()
        ms, _ = timed_region(fn, reps, 0)
        res[name] = {"ms": round(ms, 5)}
        if nbytes:
            res[name]["gbps"] = round(nbytes / (ms / 1e3) / 1e9, 1)
        log("%s: %s" % (name, res[name]))

    st = {}

    def pre():
        st["mb"] = ibp.preprocess(x, None)
    ev("preprocess", pre, x.numel() * 8)
    mask, bitval = st["mb"] if "mb" in st else ibp.preprocess(x, None)
    res["mask_checksum"] = "%x:%x" % (int(mask.view(torch.int64).sum()) & 0xffffffffffff,
                                      int(bitval.view(torch.int64).sum()) & 0xffffffffffff)
    ev("get_compress_size", lambda: ibp.get_compress_size(x, mask, bitval), x.numel() * 8)
    comp = x.clone()
    src = x.clone()

    def pack():
        comp.copy_(src)
        st["bm"] = ibp.compress_inplace(comp, mask, bitval)
    ev("copy_and_compress_inplace", pack)
    if "bm" not in st:
        st["bm"] = ibp.compress_inplace(comp, mask, bitval)
    bm = st["bm"]
    res["pack_checksum"] = "%x:%x" % (int(comp.view(-1)[::7].sum()) & 0xffffffffffff,
                                      int(bm.to(torch.int64).sum()) & 0xffffffffffff)
    cl = ibp.get_single_compress_len(x, mask, bitval)
    idx = workloads.gnn_minibatch_indices(n, min(args.batch, n), 0).cuda()
    out = torch.empty((idx.numel(), v), dtype=torch.int64, device="cuda")
    dec = out.numel() * 8
    ev("fetch_hbm", lambda: ibp.decompress_fetch(comp, mask, bitval, bm, dev, cl, idx, None, None, None, out), dec)
    ok = ops is None or "fetch_hbm" not in ops or bool(torch.equal(out, x[idx]))
    zero = torch.zeros_like(bm)
    ev("raw_gather_hbm", lambda: ibp.decompress_fetch(x, mask, bitval, zero, dev, v, idx, None, None, 2, out), dec)
    ok = ok and (ops is None or "raw_gather_hbm" not in ops or bool(torch.equal(out, x[idx])))
    ev("torch_index_gather_hbm", lambda: torch.index_select(x, 0, idx, out=out), dec)
    if want("compress_condensed") or want("fetch_condensed_hbm"):
        def cond():
            st["cc"] = ibp.compress_condensed(x, mask, bitval)
        ev("compress_condensed", cond, x.numel() * 8)
        if "cc" not in st:
            cond()
        stream, offs = st["cc"]
        ev("fetch_condensed_hbm", lambda: ibp.decompress(stream, offs, mask, bitval, dev, v, idx, out, cl), dec)
        ok = ok and (not want("fetch_condensed_hbm") or bool(torch.equal(out, x[idx])))
        del st["cc"], stream, offs
    if want("device_copy_256MiB"):
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        ev("device_copy_256MiB", lambda: b.copy_(a), 2 * a.numel())
        del a, b
    if want("seq_copy_hbm"):
        # the raw-row kernel over every row in order: a plain copy of the
        # whole table (n*V*8 bytes read and written), the calibration of the
        # HBM counters for the fetch kernel's family
        whole = torch.empty_like(x)
        ev("seq_copy_hbm", lambda: ibp.decompress_fetch(x, mask, bitval, zero, dev, v, None, None, None, 2, whole),
           2 * x.numel() * 8)
        ok = ok and bool(torch.equal(whole, x))
        del whole
    if not args.no_pcie and want("fetch_pinned"):
        comp_h = comp.cpu().pin_memory()
        ev("fetch_pinned", lambda: ibp.decompress_fetch(comp_h, mask, bitval, bm, dev, cl, idx, None, None, None,
                                                        out), dec)
        ok = ok and bool(torch.equal(out, x[idx]))
        del comp_h
    if not args.no_pcie and (want("preprocess_pinned") or want("kmeans_pinned") or want("preprocess_pinned_direct")):
        x_h = x.cpu().pin_memory()
        pm = {}

        def pre_h():
            pm["mb"] = ibp.preprocess(x_h, None)
        # the pinned table: copied to HBM once, then both passes read HBM
        ev("preprocess_pinned", pre_h, x.numel() * 8)
        if want("preprocess_pinned_direct"):
            # the C ABI on the pinned table itself: both passes read it over PCIe
            from invariantbitpacking_amd import _capi
            md = torch.zeros(v, dtype=torch.int64, device="cuda")
            bd = torch.zeros(v, dtype=torch.int64, device="cuda")
            ws_b = _capi.lib().ibp_preprocess_workspace_size(8, v, -1.0)
            ws = torch.empty(ws_b, dtype=torch.uint8, device="cuda")
            ev("preprocess_pinned_direct",
               lambda: _capi.call("ibp_preprocess", 8, x_h.data_ptr(), n, v, -1.0, md.data_ptr(), bd.data_ptr(), None,
                                  ws.data_ptr(), ws_b, torch.cuda.current_stream().cuda_stream), x.numel() * 8)
            ok = ok and bool(torch.equal(md, mask)) and bool(torch.equal(bd, bitval))
        if "mb" in pm:
            ok = ok and bool(torch.equal(pm["mb"][0], mask)) and bool(torch.equal(pm["mb"][1], bitval))
        if want("kmeans_pinned"):
            blk_h = x_h[:4096 * 8].view(4096, -1)  # 4096 rows of 400 elements (an LLM-block-like shape)
            blk_d = blk_h.cuda()
            ev("kmeans_pinned", lambda: pm.__setitem__("km", ibp.preprocess_kmeans(blk_h, 16, 0.85, seed=1)),
               blk_h.numel() * 8)
            want_km = ibp.preprocess_kmeans(blk_d, 16, 0.85, seed=1)
            if "km" in pm:
                ok = ok and all(bool(torch.equal(a.cpu(), b.cpu())) for a, b in zip(pm["km"], want_km))
        del x_h
    for k in (4, 16):
        # k-means of the whole table (beyond 2^20 rows: the many-block grouping)
        ev("kmeans_table_k%d" % k, lambda k=k: ibp.preprocess_kmeans(x, k, 0.85, seed=1), x.numel() * 8)
    if want("kmeans_llm_1024") or want("kmeans_llm_2752"):
        layer = workloads.llm_weight_blocks(layers=1)[0]
        for blk in (layer[0], layer[-1]):
            xb = blk.cuda()
            ev("kmeans_llm_%d" % xb.shape[1], lambda xb=xb: ibp.preprocess_kmeans(xb, 16, 0.85, seed=1),
               xb.numel() * 8)
        del layer, xb
    if not args.no_pcie and want("raw_gather_pinned"):
        x_h = x.cpu().pin_memory()
        ev("raw_gather_pinned", lambda: ibp.decompress_fetch(x_h, mask, bitval, zero, dev, v, idx, None, None, 2,
                                                             out), dec)
        ok = ok and bool(torch.equal(out, x[idx]))
        del x_h
    if not args.no_pcie and want("sequential_read_pinned_256MiB"):
        rows = (256 << 20) // 512
        seq_h = torch.zeros((rows, 64), dtype=torch.int64).pin_memory()
        seq_d = torch.empty((rows, 64), dtype=torch.int64, device="cuda")
        z64 = torch.zeros(64, dtype=torch.int64, device="cuda")
        zb = torch.zeros((rows + 31) // 32, dtype=torch.int32, device="cuda")
        ev("sequential_read_pinned_256MiB", lambda: ibp.decompress_fetch(seq_h, z64, z64, zb, dev, 64, None, None,
                                                                         None, 2, seq_d), rows * 512)
    res["bitexact"] = ok
    return res


def dec_bytes(nbytes, ms):
    """GB/s of nbytes moved in ms milliseconds."""
    return nbytes / (ms / 1e3) / 1e9


def run_llm(args):
    """LLM weight offload: fp16 N(0, 0.02) blocks [4096 x 4096] x4 and
    [4096 x 11008] x3 per decoder layer, viewed as int64 [4096 x 1024] /
    [4096 x 2752]; preprocess_kmeans(16) per block; compress_inplace in pinned
    host memory; layer-by-layer decompress_fetch (the kernel reads host memory
    over PCIe) while a bf16 GEMM loop runs on a second stream.  Coded twice:
    with mask row 0 for every row (what the reference's compress_inplace
    applies) and with each row's own cluster mask (clusters=)."""
    import invariantbitpacking_amd as ibp
    from invariantbitpacking_amd import workloads
    layers = workloads.llm_weight_blocks(layers=args.layers)
    blocks = []
    t_km = t_pack = 0.0
    raw = comp_b = comp_c = 0
    for li, layer in enumerate(layers):
        for w in layer:
            xd = w.cuda()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            mask, bitval, clusters = ibp.preprocess_kmeans(xd, 16, 0.85, seed=li + 1)
            torch.cuda.synchronize()
            t_km += time.perf_counter() - t0
            host = w.clone().pin_memory()
            sizes = ibp.get_compress_size(xd, mask, bitval)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            bm = ibp.compress_inplace(host, mask, bitval)
            torch.cuda.synchronize()
            t_pack += time.perf_counter() - t0
            raw += w.numel() * 8
            comp_b += int(sizes.sum())
            cl = clusters.cuda()
            comp_c += int(ibp.get_compress_size(xd, mask, bitval, clusters=cl).sum())
            host_c = w.clone().pin_memory()
            bm_c = ibp.compress_inplace(host_c, mask, bitval, clusters=cl)
            blocks.append(dict(x=w, host=host, mask=mask, bitval=bitval, bitmask=bm,
                               comp_len=ibp.get_single_compress_len(xd, mask, bitval),
                               out=torch.empty_like(xd), dev=None, host_c=host_c, bitmask_c=bm_c, clusters=cl))
            del xd
    log("llm: %d layers, %d blocks, ratio %.4f, kmeans %.2fs, pack %.2fs" % (len(layers), len(blocks),
                                                                            raw / comp_b, t_km, t_pack))
    per_layer = len(blocks) // len(layers)

    def fetch_layer(li, src="host"):
        for b in blocks[li * per_layer:(li + 1) * per_layer]:
            if src == "host_c":
                ibp.decompress_fetch(b[src], b["mask"], b["bitval"], b["bitmask_c"], torch.device("cuda"),
                                     b["comp_len"], None, None, None, None, b["out"], clusters=b["clusters"])
            else:
                ibp.decompress_fetch(b[src], b["mask"], b["bitval"], b["bitmask"], torch.device("cuda"),
                                     b["comp_len"], None, None, None, None, b["out"])

    def fetch_all(src="host"):
        for li in range(len(layers)):
            fetch_layer(li, src)
    c_ms, _ = timed_region(lambda: fetch_all("host_c"), 3, 1)
    c_exact = all(bool(torch.equal(b["out"].cpu(), b["x"])) for b in blocks)
    for b in blocks:
        del b["host_c"]
    f_ms, _ = timed_region(fetch_all, 3, 1)
    exact = all(bool(torch.equal(b["out"].cpu(), b["x"])) for b in blocks)
    # device-resident fetch of the same blocks (wide-row kernel against HBM)
    for b in blocks:
        b["dev"] = b["host"].cuda()
    d_ms, _ = timed_region(lambda: fetch_all("dev"), 5, 1)
    d_exact = all(bool(torch.equal(b["out"].cpu(), b["x"])) for b in blocks)
    # overlap: GEMM loop on a second stream while the layers are fetched
    a = torch.randn((8192, 8192), device="cuda", dtype=torch.bfloat16)
    bb = torch.randn((8192, 8192), device="cuda", dtype=torch.bfloat16).contiguous()
    side = torch.cuda.Stream()

    def gemms(k):
        for _ in range(k):
            torch.matmul(a, bb)
    # the first GEMM on a new stream sets up its BLAS workspace (milliseconds,
    # with a device synchronisation): keep that out of the overlap timing
    with torch.cuda.stream(side):
        gemms(1)
    torch.cuda.synchronize()
    g_ms, _ = timed_region(lambda: gemms(4), 3, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(side):
        gemms(int(max(1, round(f_ms / (g_ms / 4)))))
    fetch_all()
    torch.cuda.synchronize()
    both = (time.perf_counter() - t0) * 1e3
    g_alone = int(max(1, round(f_ms / (g_ms / 4)))) * g_ms / 4
    dec = raw
    return {"workload": "llm offload: %d layers x (4 x [4096x4096] + 3 x [4096x11008]) fp16 as int64 rows of "
                        "1024/2752, kmeans(16) masks (row 0 used by the codec)" % len(layers),
            "per_cluster": {"compression_ratio": round(raw / comp_c, 4),
                            "fetch_pinned_ms_per_layer": round(c_ms / len(layers), 3),
                            "fetch_pinned_decompressed_gbps": round(dec_bytes(raw, c_ms), 2),
                            "fetch_pinned_pcie_gbps": round(dec_bytes(comp_c, c_ms), 2), "bitexact": c_exact},
            "compression_ratio": round(raw / comp_b, 4), "kmeans_ms_per_block": round(t_km / len(blocks) * 1e3, 3),
            "compress_inplace_pinned_gbps": round(raw / t_pack / 1e9, 2),
            "fetch_pinned_ms_per_layer": round(f_ms / len(layers), 3),
            "fetch_pinned_decompressed_gbps": round(dec / (f_ms / 1e3) / 1e9, 2),
            "fetch_pinned_pcie_gbps": round(comp_b / (f_ms / 1e3) / 1e9, 2),
            "fetch_hbm_decompressed_gbps": round(dec / (d_ms / 1e3) / 1e9, 1),
            "fetch_hbm_alg_gbps": round((comp_b + dec) / (d_ms / 1e3) / 1e9, 1),
            "overlap": {"fetch_ms": round(f_ms, 2), "gemm_ms": round(g_alone, 2), "together_ms": round(both, 2),
                        "overlap_efficiency": round((f_ms + g_alone) / both, 3)},
            "bitexact": exact and d_exact}


def run_default(args):
    """The headline line (GNN minibatch) with the other workloads as legs:
    GNN features (4096-row calls, plain vs batched), DLRM (26 tables, plain vs
    one launch), serving (Poisson mix over the GNN features, GNN minibatch and
    DLRM sets) and LLM offload (one decoder layer, k-means masks)."""
    import copy
    traffic = None if args.no_traffic else measure_traffic(args)
    res = run_gnn_minibatch(args, traffic)
    g2 = res.pop("set")
    if args.no_legs:
        return res
    legs = {}
    f = run_gnn_features(args)
    g1 = f.pop("set")
    legs["gnn_features"] = f
    da = copy.copy(args)
    da.batch = 65536
    d = run_dlrm(da, condensed=False)
    g3 = d.pop("sets")
    legs["dlrm"] = d
    legs["serve"] = run_serve(args, groups=[[g1], [g2], g3])
    del g1, g2, g3
    torch.cuda.empty_cache()
    la = copy.copy(args)
    la.layers = 1
    legs["llm"] = run_llm(la)
    res["legs"] = legs
    res["bitexact_all_legs"] = res["bitexact"] and all(
        leg.get("bitexact", True) and all(v.get("bitexact", True) for v in leg.values() if isinstance(v, dict))
        for leg in legs.values())
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--batch", type=int, default=200000)
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the table rows (1.0 = full size)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pcie", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--workload", default="gnn_minibatch",
                    choices=["gnn_minibatch", "gnn_features", "dlrm", "serve", "llm", "pcie_rows", "kernels"],
                    help="gnn_minibatch is the headline line; the others print their own JSON")
    ap.add_argument("--rates", type=float, nargs="+", default=[1e2, 3e2, 1e3, 3e3, 1e4])
    ap.add_argument("--layers", type=int, default=32, help="decoder layers of the llm workload")
    ap.add_argument("--ops", default="", help="kernels workload: comma-separated subset of its ops")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes of the default run")
    ap.add_argument("--no-legs", action="store_true",
                    help="default run: only the GNN minibatch line, without the other workloads' legs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (rows would shard by range, "
              "each GPU fetching its share of an index batch over its own PCIe link)", file=sys.stderr)
        sys.exit(2)
    # device_count() does not initialise the GPU: the counter passes of the
    # default run start child processes before this process touches it.
    if torch.cuda.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.workload == "dlrm":
        if args.batch == 200000:
            args.batch = 65536
        r = run_dlrm(args)
        r.pop("sets")
        print(json.dumps(r))
    elif args.workload == "serve":
        print(json.dumps(run_serve(args)))
    elif args.workload == "llm":
        print(json.dumps(run_llm(args)))
    elif args.workload == "gnn_features":
        r = run_gnn_features(args)
        r.pop("set")
        print(json.dumps(r))
    elif args.workload == "pcie_rows":
        print(json.dumps(run_pcie_rows(args)))
    elif args.workload == "kernels":
        print(json.dumps(run_kernels(args)))
    else:
        print(json.dumps(run_default(args)))


if __name__ == "__main__":
    main()
