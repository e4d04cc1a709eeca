"""Adapters for the callers the reference was built for.

Each replaces the host->GPU copy of one workload with a fetch from
compressed pinned host memory, as the reference describes it:

* FlexGen weight offload -- ``CompressedHostTensor`` + ``general_copy``
  (README.md:100-143: ``ibp_preprocess``/``ibp_compress`` on the CPU copy of a
  weight, ``decompress_fetch(..., output_tensor=dst)`` in place of
  ``dst.copy_(src)``);
* ColossalAI ``CachedEmbeddingBag(comp=True)`` -- ``CompressedEmbeddingBag``
  (tests/dlrm_comp.py:79-81): the table lives compressed in pinned memory, a
  device cache holds recently used rows, misses are fetched by index;
* DGL / Legion feature loading -- ``FeatureStore``: minibatch node features
  gathered from compressed pinned features into a reusable device buffer.

Values are bit-identical to the uncompressed path: IBP is lossless and every
fetch goes through the HIP kernels of libibp.
"""
import torch
import torch.nn.functional as F

from . import compress_inplace, decompress_fetch, get_single_compress_len, preprocess, preprocess_kmeans

_WORD = {8: torch.int64, 4: torch.int32, 2: torch.int16, 1: torch.int8}


def _words(t):
    """Widest integer view of a contiguous tensor's bytes along its last dim
    (int64 where the row length allows: fewer header bits per row)."""
    row_bytes = t.size(-1) * t.element_size()
    w = next(w for w in (8, 4, 2, 1) if row_bytes % w == 0)
    return t.view(_WORD[w])


class CompressedHostTensor:
    """A 2-D tensor compressed in place in pinned host memory.

    ``clusters``: 0 uses one mask for all rows (``preprocess``, as FlexGen
    does in the README); K > 0 runs ``preprocess_kmeans(K)`` and codes every
    row with its own cluster's mask.
    """

    def __init__(self, tensor, threshold=None, clusters=0, seed=1):
        assert tensor.dim() == 2, "2-D tensors only (view larger ones as 2-D)"
        self.dtype, self.shape = tensor.dtype, tuple(tensor.shape)
        host = tensor.detach().contiguous().cpu()
        self.data = host.clone().pin_memory() if not host.is_pinned() else host.clone()
        words = _words(self.data)
        dev = words.cuda()
        self.clusters = None
        if clusters:
            self.mask, self.bitval, cl = preprocess_kmeans(dev, int(clusters), threshold, seed=seed)
            self.clusters = cl.cuda()
        else:
            self.mask, self.bitval = preprocess(dev, threshold)
        # comp_len is a read-size hint, computed on the uncompressed rows
        self.comp_len = get_single_compress_len(dev, self.mask[0] if self.mask.dim() == 2 else self.mask,
                                                self.bitval[0] if self.bitval.dim() == 2 else self.bitval)
        del dev
        self.bitmask = compress_inplace(words, self.mask, self.bitval, clusters=self.clusters)
        self.words = words

    def fetch_into(self, dst, rows=None):
        """Decompress rows (a slice or index tensor; all rows if None) into the
        device tensor dst (same dtype, contiguous)."""
        assert dst.dtype == self.dtype and dst.is_cuda and dst.is_contiguous()
        out = _words(dst)
        index = None
        if rows is not None:
            if isinstance(rows, slice):
                start, stop, step = rows.indices(self.shape[0])
                index = torch.arange(start, stop, step, device="cuda")
            else:
                index = rows.cuda()
        return decompress_fetch(self.words, self.mask, self.bitval, self.bitmask, dst.device, self.comp_len,
                                index, None, None, None, out, clusters=self.clusters)

    def to_device(self):
        dst = torch.empty(self.shape, dtype=self.dtype, device="cuda")
        self.fetch_into(dst)
        return dst


def general_copy(dst, dst_indices, src, src_indices):
    """FlexGen's general_copy (README.md:133-143) for a compressed source:
    copies src[src_indices] into dst[dst_indices].  src is a
    CompressedHostTensor; indices select along the first dimension (a tuple
    whose first entry is a slice, or None for everything), and the
    destination region must be contiguous.  Other sources fall back to a plain
    copy, as in FlexGen."""
    if not isinstance(src, CompressedHostTensor):
        s = src[src_indices] if src_indices else src
        d = dst[dst_indices] if dst_indices else dst
        d.copy_(s, non_blocking=True)
        return
    rows = None
    if src_indices:
        rows = src_indices[0]
        assert all(i == slice(None) for i in src_indices[1:]), "row slices only"
    d = dst[dst_indices] if dst_indices else dst
    assert d.is_contiguous(), "the destination region must be contiguous"
    src.fetch_into(d, rows)


class CompressedEmbeddingBag(torch.nn.Module):
    """Embedding bag whose table stays compressed in pinned host memory.

    A device cache of ``cache_rows`` rows serves hits; the rows of a batch
    that miss are fetched from the compressed table in one decompress_fetch
    (ColossalAI CachedEmbeddingBag with comp=True, tests/dlrm_comp.py:79-81).
    Eviction is FIFO over cache slots, skipping slots that hold rows of the
    current batch.  Output equals
    ``F.embedding_bag(input, weight, offsets, mode=mode)`` on the original
    weight.
    """

    def __init__(self, weight, cache_rows, mode="sum", threshold=None):
        super().__init__()
        assert weight.dim() == 2
        self.num_rows, self.dim = weight.shape
        self.mode = mode
        self.table = CompressedHostTensor(weight, threshold)
        self.cache_rows = int(cache_rows)
        self.cache = torch.zeros((self.cache_rows, self.dim), dtype=weight.dtype, device="cuda")
        self.slot_of = torch.full((self.num_rows,), -1, dtype=torch.int64, device="cuda")  # row -> slot
        self.row_in = torch.full((self.cache_rows,), -1, dtype=torch.int64, device="cuda")  # slot -> row
        self.next_slot = 0
        self.misses = 0
        self.lookups = 0

    def _admit(self, ids):
        """Puts the unique rows `ids` (device int64) in the cache; returns their slots."""
        assert ids.numel() <= self.cache_rows, "a batch touches more rows than the cache holds"
        slots = self.slot_of[ids]
        hit = slots >= 0
        miss = ids[~hit]
        m = int(miss.numel())
        self.lookups += int(ids.numel())
        self.misses += m
        if m:
            # victims: the next m slots in FIFO order that hold no row of this batch
            order = (torch.arange(self.cache_rows, device="cuda") + self.next_slot) % self.cache_rows
            busy = torch.zeros(self.cache_rows, dtype=torch.bool, device="cuda")
            busy[slots[hit]] = True
            new = order[~busy[order]][:m]
            self.next_slot = (int(new[-1]) + 1) % self.cache_rows
            old = self.row_in[new]
            self.slot_of[old[old >= 0]] = -1  # evict
            stage = torch.empty((m, self.dim), dtype=self.cache.dtype, device="cuda")
            self.table.fetch_into(stage, miss)
            self.cache[new] = stage
            self.row_in[new] = miss
            self.slot_of[miss] = new
            slots = self.slot_of[ids]
        return slots

    def forward(self, input, offsets=None, per_sample_weights=None):
        ids, inverse = torch.unique(input, return_inverse=True)
        slots = self._admit(ids)
        return F.embedding_bag(slots[inverse], self.cache, offsets, mode=self.mode,
                               per_sample_weights=per_sample_weights)


class FeatureStore:
    """Node features compressed in pinned host memory; ``gather(nodes)``
    decompresses a minibatch's rows into a reusable device buffer (the
    DGL / Legion feature-loading step)."""

    def __init__(self, features, threshold=None):
        self.table = CompressedHostTensor(features, threshold)
        self._buf = None

    def gather(self, nodes, out=None):
        n = int(nodes.numel())
        if out is None:
            if self._buf is None or self._buf.size(0) < n:
                self._buf = torch.empty((max(n, 1), self.table.shape[1]), dtype=self.table.dtype, device="cuda")
            out = self._buf[:n]
        self.table.fetch_into(out, nodes)
        return out
