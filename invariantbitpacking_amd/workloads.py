"""Seeded synthetic data of the benchmark workloads (no datasets are downloaded).

Each generator returns CPU tensors; `scale` < 1 shrinks row counts for quick
tests while keeping the row shape and value distribution.

  gnn_features     -- fp32 [65536 x 128] ~ N(0, 0.1), 30% exact zeros, viewed
                      as int64 [65536 x 64]; 4096 uniform fetch indices.
  gnn_minibatch    -- fp32 [2,449,029 x 100] ~ N(0, 0.5) clipped to [-4, 4],
                      25% zeros, 10% of columns constant within 4096-row
                      blocks, viewed as int64 [rows x 50]; minibatches of
                      200,000 uniform indices.
  dlrm_tables      -- 26 fp32 tables of dim 128, rows log-uniform in
                      [1e3, 1e7], N(0, 0.01) quantised to 2^-12, viewed as
                      int64 [rows x 64]; 65,536 Zipf(1.05) rows per table.
  llm_weight_blocks-- fp16 N(0, 0.02) blocks [4096 x 4096] and [4096 x 11008]
                      viewed as int64 [4096 x 1024] / [4096 x 2752].
"""
import math

import numpy as np
import torch


def _rows(n, scale):
    return max(1, int(round(n * scale)))


def gnn_features(scale=1.0, seed=0):
    n = _rows(65536, scale)
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, 0.1, size=(n, 128)).astype(np.float32)
    x[rng.random(x.shape) < 0.3] = 0.0
    data = torch.from_numpy(x).view(torch.int64)
    idx = torch.from_numpy(np.random.default_rng(1).integers(0, n, size=min(4096, n * 4), dtype=np.int64))
    return data, idx


def gnn_minibatch(scale=1.0, seed=0, device="cpu", block_rows=4096):
    """Generated with torch on `device` (fast on the GPU) and returned there."""
    n = _rows(2449029, scale)
    g = torch.Generator(device=device).manual_seed(seed)
    x = torch.randn((n, 100), generator=g, device=device, dtype=torch.float32).mul_(0.5).clamp_(-4.0, 4.0)
    x.masked_fill_(torch.rand((n, 100), generator=g, device=device) < 0.25, 0.0)
    cols = torch.randperm(100, generator=g, device=device)[:10]
    nblocks = (n + block_rows - 1) // block_rows
    consts = torch.randn((nblocks, 10), generator=g, device=device).mul_(0.5).clamp_(-4.0, 4.0)
    rep = consts.repeat_interleave(block_rows, dim=0)[:n]
    x[:, cols] = rep
    return x.view(torch.int64)


def gnn_minibatch_indices(num_rows, batch=200000, step=0):
    g = torch.Generator().manual_seed(2 + step)
    return torch.randint(0, num_rows, (batch,), generator=g, dtype=torch.int64)


def dlrm_table_rows(num_tables=26, scale=1.0, seed=100):
    """Row counts of the tables: log-uniform in [1e3, 1e7] (seeded), times scale."""
    rng = np.random.default_rng(seed)
    rows = np.round(10 ** rng.uniform(3, 7, size=num_tables)).astype(np.int64)
    return [max(64, int(round(r * scale))) for r in rows]


def dlrm_tables(num_tables=26, scale=1.0, seed=100, device="cpu"):
    """fp32 [rows x 128] ~ N(0, 0.01) quantised to 2^-12, viewed as int64 [rows x 64]."""
    tables = []
    for t, r in enumerate(dlrm_table_rows(num_tables, scale, seed)):
        g = torch.Generator(device=device).manual_seed(seed + 1000 + t)
        x = torch.randn((r, 128), generator=g, device=device, dtype=torch.float32).mul_(0.01)
        x = torch.round(x * 4096.0).div_(4096.0)
        tables.append(x.view(torch.int64))
    return tables


def dlrm_indices(table_rows, batch=65536, seed=101):
    out = []
    for t, r in enumerate(table_rows):
        g = np.random.default_rng(seed + t)
        z = g.zipf(1.05, size=batch) - 1
        bad = z >= r
        while bad.any():
            z[bad] = g.zipf(1.05, size=int(bad.sum())) - 1
            bad = z >= r
        out.append(torch.from_numpy(z.astype(np.int64)))
    return out


def llm_weight_blocks(layers=1, scale=1.0, seed=200, device="cuda"):
    """Per layer: 4 x [4096 x 4096] and 3 x [4096 x 11008] fp16 N(0, 0.02)
    blocks as int64 views, generated on `device` and returned on the CPU."""
    if device == "cuda" and not torch.cuda.is_available():
        device = "cpu"
    out = []
    for layer in range(layers):
        g = torch.Generator(device=device).manual_seed(seed + layer)
        rows = max(4, int(4096 * scale))
        blocks = []
        for shape in [(rows, 4096)] * 4 + [(rows, 11008)] * 3:
            w = (torch.randn(shape, generator=g, device=device, dtype=torch.float32) * 0.02).to(torch.float16)
            blocks.append(w.view(torch.int64).cpu())
        out.append(blocks)
    return out


def log_uniform_int(rng, lo, hi, size=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=size)).astype(np.int64)
