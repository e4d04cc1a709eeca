"""Batched fetch: many tables fetched and unpacked by ONE kernel launch.

No reference counterpart (the reference fetches one table per call); each
table's result is exactly what decompress_fetch returns for it.  The plan
captures pointers, so index and output buffers are refilled in place between
launches; `capture_graph()` records the launch into a HIP graph.
"""
import ctypes
import hashlib
import json
import os

import torch

from . import _capi


class BatchedFetch:
    """tasks: list of dicts with keys comp, mask, bitval, bitmask, out and
    optionally index (int32/int64 GPU tensor), comp_len (int).

    Condensed streams (compress_condensed): comp is the 1-D stream, with
    offsets (its int64 byte offsets, GPU or pinned) and vec_size in place of
    bitmask; a plan's tasks are then all condensed, with rows of <= 64
    elements of one element width and one index width."""

    def __init__(self, tasks):
        self._keep = []
        arr = (_capi.FetchTask * len(tasks))()
        for i, t in enumerate(tasks):
            comp, out = t["comp"], t["out"]
            index = t.get("index")
            offsets = t.get("offsets")
            if offsets is not None:
                if offsets.dtype != torch.int64 or offsets.dim() != 1 or not offsets.is_contiguous():
                    raise RuntimeError("task %d: offsets must be a contiguous 1-D int64 tensor" % i)
                if not (offsets.is_cuda or offsets.is_pinned()):
                    raise RuntimeError("task %d: offsets must be on the GPU or pinned" % i)
                if comp.dim() != 1 or not comp.is_contiguous():
                    raise RuntimeError("task %d: a condensed stream must be a contiguous 1-D tensor" % i)
                num_vecs, v = offsets.numel(), int(t["vec_size"])
            else:
                num_vecs, v = (1, comp.size(0)) if comp.dim() == 1 else tuple(comp.shape)
            n = num_vecs if index is None else index.numel()
            if out.shape != (n, v) or out.dtype != comp.dtype or not out.is_contiguous():
                raise RuntimeError("task %d: output must be a contiguous [%d, %d] %s tensor" % (i, n, v, comp.dtype))
            if index is not None:
                if not index.is_cuda or index.dim() != 1:
                    raise RuntimeError("task %d: index must be a 1-D GPU tensor" % i)
                if index.dtype not in (torch.int32, torch.int64):
                    index = index.to(torch.int64)
                index = index.contiguous()
            bitmask = t.get("bitmask")
            self._keep += [comp, t["mask"], t["bitval"], bitmask, out, index, offsets]
            arr[i] = _capi.FetchTask(
                comp.data_ptr(), t["mask"].data_ptr(), t["bitval"].data_ptr(),
                None if bitmask is None else bitmask.data_ptr(),
                None if index is None else index.data_ptr(), out.data_ptr(), num_vecs, v, n,
                comp.element_size(), 0 if index is None else index.element_size(), int(t.get("comp_len") or 0),
                1 if comp.is_cuda else 2, None if offsets is None else offsets.data_ptr(),
                0 if offsets is None else comp.numel() * comp.element_size())
        self._tasks = arr
        self._signature = [[int(t.elem_bytes), int(t.vec_size), int(t.n), int(t.num_vecs), int(t.index_bytes),
                            int(t.comp_loc), int(t.comp_len), 1 if t.offsets else 0] for t in arr]
        self._plan = ctypes.c_void_p()
        _capi.call("ibp_plan_create", arr, len(tasks), ctypes.byref(self._plan))
        self.tasks = tasks

    @staticmethod
    def _stream():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def launch(self):
        _capi.call("ibp_plan_launch", self._plan, self._stream())

    def autotune(self, cache=None):
        """Times the candidate launch configurations and keeps the fastest;
        returns its milliseconds per launch.  cache: path of a JSON file of
        earlier results, keyed by the GPU and the plan's task shapes
        (element and index widths, row length, rows, where the rows live,
        comp_len); a hit applies the stored configuration without timing
        anything, a miss tunes and adds the result."""
        key = None
        if cache is not None:
            key = self._cache_key()
            stored = self._cache_read(cache).get(key)
            if stored is not None:
                self.set_config(stored["config"])
                return float(stored["ms"])
        ms = ctypes.c_float(0)
        _capi.call("ibp_plan_autotune", self._plan, self._stream(), ctypes.byref(ms))
        if cache is not None:
            entries = self._cache_read(cache)
            entries[key] = {"config": self.config(), "ms": round(float(ms.value), 5), "tasks": len(self._signature),
                            "device": torch.cuda.get_device_name()}
            tmp = "%s.tmp%d" % (cache, os.getpid())
            with open(tmp, "w") as f:
                json.dump(entries, f, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return ms.value

    def config(self):
        """The plan's launch configuration (see ibp_plan_get_config)."""
        cfg = (ctypes.c_int32 * 7)()
        _capi.call("ibp_plan_get_config", self._plan, cfg)
        return [int(x) for x in cfg]

    def set_config(self, cfg):
        """Applies a configuration from config() or an autotune cache."""
        arr = (ctypes.c_int32 * 7)(*[int(x) for x in cfg])
        _capi.call("ibp_plan_set_config", self._plan, arr)

    def _cache_key(self):
        blob = json.dumps([torch.cuda.get_device_name(), self._signature], separators=(",", ":"))
        return hashlib.sha1(blob.encode()).hexdigest()

    @staticmethod
    def _cache_read(path):
        try:
            with open(path) as f:
                data = json.load(f)
            return data if isinstance(data, dict) else {}
        except (OSError, ValueError):
            return {}

    def capture_graph(self):
        _capi.call("ibp_plan_capture_graph", self._plan, self._stream())

    def graph_launch(self):
        _capi.call("ibp_plan_graph_launch", self._plan, self._stream())

    def describe(self):
        buf = ctypes.create_string_buffer(512)
        _capi.call("ibp_plan_describe", self._plan, buf, 512)
        return buf.value.decode()

    def close(self):
        if self._plan:
            _capi.lib().ibp_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
