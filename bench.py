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
fetch against a measured pinned H2D copy ceiling, and p50/p99 call latency.
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


def timed_calls(fn, steps, warmup):
    """Runs fn() warmup + steps times on the current stream; returns per-call
    milliseconds (HIP events around each call) and the wall time of the timed
    region (synchronised on both sides)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in ev:
        s.record()
        fn()
        e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return [s.elapsed_time(e) for s, e in ev], wall


def h2d_ceiling_gbps(nbytes=256 << 20, reps=5):
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


def algorithmic_hbm_bytes(sizes_sel, n, v, w):
    """Per launch: compressed bytes read + decompressed bytes written + int64
    index + bitmask words (ceil(n/32)*4)."""
    return int(sizes_sel) + n * v * w + n * 8 + (n + 31) // 32 * 4


def run_gnn_minibatch(args):
    import invariantbitpacking_amd as ibp
    from invariantbitpacking_amd import workloads

    t0 = time.perf_counter()
    x = workloads.gnn_minibatch(scale=args.scale, device="cuda")
    num_rows, v = x.shape
    w = 8
    log("data %s generated in %.1fs" % (tuple(x.shape), time.perf_counter() - t0))
    mask, bitval = ibp.preprocess(x, None)
    sizes = ibp.get_compress_size(x, mask, bitval)
    comp = x.clone()
    bitmask = ibp.compress_inplace(comp, mask, bitval)
    comp_len = ibp.get_single_compress_len(x, mask, bitval)
    torch.cuda.synchronize()
    packed = int(torch.count_nonzero((sizes < v * w)))
    log("compressed: %d/%d rows packed, comp_len %d, ratio %.4f" %
        (packed, num_rows, comp_len, num_rows * v * w / float(sizes.sum())))

    batch = min(args.batch, num_rows)
    nb = max(args.steps + args.warmup, 1)
    idx_cpu = [workloads.gnn_minibatch_indices(num_rows, batch, s % 64) for s in range(min(nb, 64))]
    idx = [i.cuda() for i in idx_cpu]
    out = torch.empty((batch, v), dtype=torch.int64, device="cuda")
    step = [0]

    def fetch():
        ibp.decompress_fetch(comp, mask, bitval, bitmask, torch.device("cuda"), comp_len,
                             idx[step[0] % len(idx)], None, None, None, out)
        step[0] += 1

    # -- plain per-call fetch, compressed data resident in HBM (the metric)
    ms, wall = timed_calls(fetch, args.steps, args.warmup)
    last = idx[(step[0] - 1) % len(idx)]
    bitexact = bool(torch.equal(out, x[last]))
    dec_bytes = batch * v * w
    value = args.steps * dec_bytes / wall / 1e9
    kern_ms = float(np.mean(ms))
    sel = [int(sizes[i].sum()) for i in idx]
    traffic_alg = algorithmic_hbm_bytes(np.mean(sel), batch, v, w)
    achieved = traffic_alg / (kern_ms / 1e3) / 1e9
    log("plain: %.3f ms/call (events), %.1f GB/s decompressed, kernel-side %.1f GB/s HBM" % (kern_ms, value, achieved))

    # -- batched: every minibatch of the step list in one persistent launch
    outs = [torch.empty((batch, v), dtype=torch.int64, device="cuda") for _ in idx]
    bf = ibp.BatchedFetch([dict(comp=comp, mask=mask, bitval=bitval, bitmask=bitmask, index=i, out=o,
                                comp_len=comp_len) for i, o in zip(idx, outs)])
    bf.autotune()
    bf.capture_graph()
    bms, bwall = timed_calls(bf.graph_launch, max(3, args.steps // len(idx)), 1)
    batched = len(idx) * dec_bytes / (float(np.mean(bms)) / 1e3) / 1e9
    b_exact = all(bool(torch.equal(o, x[i])) for o, i in zip(outs, idx))
    log("batched: %d minibatches/launch %.3f ms, %.1f GB/s (%s)" % (len(idx), float(np.mean(bms)), batched,
                                                                    bf.describe()))
    bf.close()
    del outs

    # -- PCIe-resident: compressed data in pinned host memory
    pcie = None
    if not args.no_pcie:
        comp_h = comp.cpu().pin_memory()
        ceiling = h2d_ceiling_gbps()

        def fetch_h():
            ibp.decompress_fetch(comp_h, mask, bitval, bitmask, torch.device("cuda"), comp_len,
                                 idx[step[0] % len(idx)], None, None, None, out)
            step[0] += 1
        hms, hwall = timed_calls(fetch_h, max(3, args.steps // 2), 1)
        hk = float(np.mean(hms))
        pcie_bytes = float(np.mean(sel))
        pcie = {"decompressed_gbps": round(dec_bytes / (hk / 1e3) / 1e9, 2),
                "compressed_pcie_gbps": round(pcie_bytes / (hk / 1e3) / 1e9, 2),
                "h2d_copy_ceiling_gbps": round(ceiling, 2),
                "frac_of_ceiling": round(pcie_bytes / (hk / 1e3) / 1e9 / ceiling, 3),
                "ms_per_call": round(hk, 4),
                "bitexact": bool(torch.equal(out, x[idx[(step[0] - 1) % len(idx)]]))}
        log("pcie: %s" % pcie)
        del comp_h

    # -- CPU baseline (oracle, one core)
    cpu = None
    if not args.no_cpu:
        comp_np = comp.cpu().numpy()
        om = mask.cpu().numpy().view(np.uint64)
        ob = bitval.cpu().numpy().view(np.uint64)
        obm = bitmask.cpu().numpy()
        cpu = cpu_baseline(comp_np, om, ob, obm, [i.numpy() for i in idx_cpu], budget_s=args.cpu_seconds)
        log("cpu baseline: %s" % cpu)

    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": "gnn_minibatch: %d x %d int64 (fp32 [%d x 100]) compressed in HBM, %d random rows "
                               "per step, decompress_fetch into a preallocated output" % (num_rows, v, num_rows, batch),
                   "rows": num_rows, "vec_size": v, "batch": batch, "packed_rows": packed, "comp_len": comp_len},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": traffic_alg, "kernel_ms": round(kern_ms, 5)},
        "cpu_baseline": cpu,
        "bitexact": bitexact and b_exact,
        "latency_ms": {"p50": round(float(np.percentile(ms, 50)), 4), "p99": round(float(np.percentile(ms, 99)), 4)},
        "batched": {"decompressed_gbps": round(batched, 2), "minibatches_per_launch": len(idx),
                    "vs_plain": round(batched / value, 3)},
        "pcie_resident": pcie,
    }
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
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(rows shard per GPU over each GPU's own PCIe link; see the design notes)", file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    res = run_gnn_minibatch(args)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
