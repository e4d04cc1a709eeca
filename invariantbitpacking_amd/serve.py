"""Served fetch: a continuous-batching request loop over resident tensor sets.

No reference counterpart.  Requests (a registered compressed tensor set plus
host row indices) are admitted when their output fits the server's device
arena, merged with whatever else is admitted into one batched kernel launch
and run on one of several streams, so requests that arrive while others are on
the GPU overlap with them.  Results are bit-identical to decompress_fetch.
"""
import ctypes

import numpy as np
import torch

from . import _capi


class _ArenaView:
    """__cuda_array_interface__ over a request's output in the arena."""

    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {"shape": shape, "typestr": typestr, "data": (ptr, False), "version": 2,
                                         "strides": None}


class Server:
    def __init__(self, hbm_budget_bytes=1 << 30, num_streams=4, max_batch_requests=64, max_batch_rows=1 << 20):
        cfg = _capi.ServerConfig(hbm_budget_bytes, num_streams, max_batch_requests, max_batch_rows)
        self._srv = ctypes.c_void_p()
        _capi.call("ibp_server_create", ctypes.byref(cfg), ctypes.byref(self._srv))
        self._sets = {}
        self._keep = []

    def register(self, comp, mask, bitval, bitmask, comp_len=None):
        """Registers a compressed tensor set (kept resident); returns its id."""
        if not (comp.is_cuda or comp.is_pinned()):
            raise RuntimeError("compressed data must be on the GPU or pinned")
        num_vecs, v = (1, comp.size(0)) if comp.dim() == 1 else tuple(comp.shape)
        sid = ctypes.c_int(0)
        _capi.call("ibp_server_register", self._srv, comp.element_size(), comp.data_ptr(), num_vecs, v,
                   mask.data_ptr(), bitval.data_ptr(), bitmask.data_ptr(), int(comp_len or 0),
                   1 if comp.is_cuda else 2, ctypes.byref(sid))
        self._keep += [comp, mask, bitval, bitmask]
        self._sets[sid.value] = (v, comp.dtype)
        return sid.value

    @staticmethod
    def clock_us():
        """The server's clock (host steady-clock microseconds)."""
        return _capi.lib().ibp_server_clock_us()

    def submit(self, set_id, index, arrival_us=None):
        """index: host int64 array / CPU tensor of row ids; returns a request id.
        arrival_us (clock_us() time, not in the future): when the request
        arrived, e.g. its scheduled arrival; latency counts from it (default:
        now)."""
        idx = np.ascontiguousarray(np.asarray(index.cpu() if torch.is_tensor(index) else index, dtype=np.int64))
        rid = ctypes.c_int64(0)
        _capi.call("ibp_server_submit_at", self._srv, set_id, idx.ctypes.data, idx.size,
                   0.0 if arrival_us is None else float(arrival_us), ctypes.byref(rid))
        return rid.value, set_id, idx.size

    def wait(self, req, timeout_ms=-1, view=True):
        """Returns (view of the [n x V] result in the arena, latency in us).  The
        view stays valid until release(req).  view=False returns (None, latency)
        without building the tensor."""
        rid, set_id, n = req
        out = ctypes.c_void_p()
        lat = ctypes.c_double(0)
        _capi.call("ibp_server_wait", self._srv, rid, int(timeout_ms), ctypes.byref(out), ctypes.byref(lat))
        if not view:
            return None, lat.value
        v, dtype = self._sets[set_id]
        typestr = {torch.int64: "<i8", torch.float64: "<f8", torch.int32: "<i4", torch.float32: "<f4"}[dtype]
        t = torch.as_tensor(_ArenaView(out.value, (n, v), typestr), device="cuda")
        return t, lat.value

    def request_times(self, req):
        """(arrival, admitted, launch start, launch issued, done) in host microseconds."""
        t = (ctypes.c_double * 5)()
        _capi.call("ibp_server_request_times", self._srv, req[0], t)
        return tuple(t)

    def release(self, req, stream=None):
        """Gives the request's output region back to the server.  The region is
        reused only after the work queued so far on `stream` (default: the
        current stream) has finished, so reads of the view queued before the
        release stay valid.  Work on other streams that reads the view must be
        finished (or ordered before `stream`) by the caller."""
        s = torch.cuda.current_stream() if stream is None else stream
        _capi.call("ibp_server_release_after", self._srv, req[0], ctypes.c_void_p(s.cuda_stream))

    def stats(self):
        st = _capi.ServerStats()
        _capi.call("ibp_server_stats_get", self._srv, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in _capi.ServerStats._fields_}

    def set_budget_override(self, nbytes):
        _capi.call("ibp_server_set_budget_override", self._srv, int(nbytes))

    def close(self):
        if self._srv:
            _capi.lib().ibp_server_destroy(self._srv)
            self._srv = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
