"""ctypes front end of the CPU oracle (libibp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.

Arrays are numpy arrays; element width w = itemsize (1, 2, 4 or 8 bytes; the
slot codec takes 4 or 8).  Signed and unsigned views are interchangeable: the
oracle works on the raw bits.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libibp_oracle.so")
_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64


def build(force=False):
    """Compile the oracle with gcc (oracle/Makefile)."""
    if force or not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_count_bits.argtypes = [ctypes.c_int, _P, _I64, _I64, _P]
        L.oracle_create_mask.argtypes = [ctypes.c_int, _P, _I64, _I64, ctypes.c_float, _P, _P]
        L.oracle_check_feats.argtypes = [ctypes.c_int, _P, _I64, _I64, _P, _P]
        L.oracle_check_feats.restype = ctypes.c_uint64
        L.oracle_preprocess.argtypes = [ctypes.c_int, _P, _I64, _I64, ctypes.c_float, _P, _P,
                                        ctypes.POINTER(ctypes.c_float)]
        L.oracle_preprocess.restype = ctypes.c_int
        L.oracle_row_size.argtypes = [ctypes.c_int, _P, _I64, _P, _P]
        L.oracle_row_size.restype = ctypes.c_int64
        L.oracle_get_compress_size.argtypes = [ctypes.c_int, _P, _I64, _I64, _P, _P, _P, _I64, _P,
                                               ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_pack_row.argtypes = [ctypes.c_int, _P, _I64, _P, _P, _P]
        L.oracle_unpack_row.argtypes = [ctypes.c_int, _P, _I64, _P, _P, _P]
        L.oracle_compress.argtypes = [ctypes.c_int, _P, _P, _I64, _I64, _P, _P, _P, _I64, _P]
        L.oracle_decompress_fetch.argtypes = [ctypes.c_int, _P, _I64, _I64, _P, _P, _P, _P, _I64, _P]
        L.oracle_get_compress_size_clustered.argtypes = [ctypes.c_int, _P, _I64, _I64, _P, _P, _P, _P, _I64,
                                                         _P, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_compress_clustered.argtypes = [ctypes.c_int, _P, _P, _I64, _I64, _P, _P, _P, _P, _I64, _P]
        L.oracle_decompress_fetch_clustered.argtypes = [ctypes.c_int, _P, _I64, _I64, _P, _P, _P, _P, _P, _I64,
                                                        _P]
        L.oracle_compress_condensed.argtypes = [ctypes.c_int, _P, _I64, _I64, _P, _P, _P, _P, _I64, _P, _P]
        L.oracle_compress_condensed.restype = ctypes.c_int64
        L.oracle_decompress_condensed.argtypes = [ctypes.c_int, _P, _P, _I64, _I64, _P, _P, _P, _P, _I64, _P]
        L.oracle_compress_mt.argtypes = [ctypes.c_int, _P, _P, _I64, _I64, _P, _P, _P, _P, _I64, _P,
                                         ctypes.c_int]
        L.oracle_random_shuffle.argtypes = [_P, _I64]
        L.oracle_kmeans.argtypes = [ctypes.c_int, _P, _I64, _I64, ctypes.c_int, _I64, _P, _P, _P]
        L.oracle_sweep_last.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.oracle_sweep_last.restype = ctypes.c_float
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _c(a):
    return np.ascontiguousarray(a)


def _uview(a):
    """Unsigned integer view of the same bits (u1/u2/u4/u8)."""
    return a.view(np.dtype("u%d" % a.itemsize))


def count_bits(data):
    data = _c(data)
    n, v = data.shape
    cnt = np.zeros(v * data.itemsize * 8, dtype=np.int32)
    lib().oracle_count_bits(data.itemsize, _ptr(data), n, v, _ptr(cnt))
    return cnt


def create_mask(cnt, num_vecs, vec_size, w, threshold):
    dt = np.dtype("u%d" % w)
    mask = np.zeros(vec_size, dtype=dt)
    bitval = np.zeros(vec_size, dtype=dt)
    cnt = _c(cnt.astype(np.int32))
    lib().oracle_create_mask(w, _ptr(cnt), num_vecs, vec_size, float(threshold), _ptr(mask), _ptr(bitval))
    return mask, bitval


def check_feats(data, mask, bitval):
    data = _c(data)
    n, v = data.shape
    return int(lib().oracle_check_feats(data.itemsize, _ptr(data), n, v,
                                        _ptr(_c(mask)), _ptr(_c(bitval))))


def preprocess(data, threshold=None):
    """Returns (mask, bitval, chosen_step, chosen_threshold)."""
    data = _c(data)
    n, v = data.shape
    dt = np.dtype("u%d" % data.itemsize)
    mask = np.zeros(v, dtype=dt)
    bitval = np.zeros(v, dtype=dt)
    chosen_t = ctypes.c_float(float("nan"))
    t = -1.0 if threshold is None else float(threshold)
    step = lib().oracle_preprocess(data.itemsize, _ptr(data), n, v, t, _ptr(mask), _ptr(bitval),
                                   ctypes.byref(chosen_t))
    return mask, bitval, step, chosen_t.value


def row_size(row, mask, bitval):
    row = _c(row)
    return int(lib().oracle_row_size(row.itemsize, _ptr(row), row.shape[-1],
                                     _ptr(_c(mask)), _ptr(_c(bitval))))


def _clusters(clusters):
    return None if clusters is None else _c(np.asarray(clusters, dtype=np.int32))


def get_compress_size(data, mask, bitval, index=None, ctr=None, clusters=None):
    """Returns (sizes int64[n], ctr) with ctr the counter after the addition.
    clusters: mask row of each source row (mask/bitval [K x V])."""
    data = _c(data)
    n_rows, v = data.shape
    idx = None if index is None else _c(np.asarray(index, dtype=np.int64))
    n = n_rows if idx is None else idx.shape[0]
    sizes = np.zeros(n, dtype=np.int64)
    c = None if ctr is None else ctypes.c_uint64(int(ctr))
    cl = _clusters(clusters)
    lib().oracle_get_compress_size_clustered(data.itemsize, _ptr(data), n_rows, v, _ptr(_c(mask)),
                                             _ptr(_c(bitval)), _ptr(cl), _ptr(idx), n, _ptr(sizes),
                                             None if c is None else ctypes.byref(c))
    return sizes, (None if c is None else c.value)


def pack_row(row, mask, bitval):
    row = _c(row)
    slot = np.zeros_like(row)
    lib().oracle_pack_row(row.itemsize, _ptr(row), row.shape[-1], _ptr(_c(mask)), _ptr(_c(bitval)), _ptr(slot))
    return slot


def unpack_row(slot, mask, bitval):
    slot = _c(slot)
    out = np.zeros_like(slot)
    lib().oracle_unpack_row(slot.itemsize, _ptr(slot), slot.shape[-1], _ptr(_c(mask)), _ptr(_c(bitval)), _ptr(out))
    return out


def compress(data, mask, bitval, index=None, clusters=None):
    """Out-of-place compress_inplace: returns (slots [n x V], bitmask int32)."""
    data = _c(data)
    n_rows, v = data.shape
    idx = None if index is None else _c(np.asarray(index, dtype=np.int64))
    n = n_rows if idx is None else idx.shape[0]
    out = np.zeros((n, v), dtype=data.dtype)
    bitmask = np.zeros((n + 31) // 32, dtype=np.int32)
    lib().oracle_compress_clustered(data.itemsize, _ptr(data), _ptr(out), n_rows, v, _ptr(_c(mask)),
                                    _ptr(_c(bitval)), _ptr(_clusters(clusters)), _ptr(idx), n, _ptr(bitmask))
    return out, bitmask


def compress_mt(data, mask, bitval, threads, index=None, clusters=None):
    """compress() on `threads` OpenMP threads (same result)."""
    data = _c(data)
    n_rows, v = data.shape
    idx = None if index is None else _c(np.asarray(index, dtype=np.int64))
    n = n_rows if idx is None else idx.shape[0]
    out = np.zeros((n, v), dtype=data.dtype)
    bitmask = np.zeros((n + 31) // 32, dtype=np.int32)
    lib().oracle_compress_mt(data.itemsize, _ptr(data), _ptr(out), n_rows, v, _ptr(_c(mask)), _ptr(_c(bitval)),
                             _ptr(_clusters(clusters)), _ptr(idx), n, _ptr(bitmask), int(threads))
    return out, bitmask


def compress_inplace(data, mask, bitval):
    """In-place on a C-contiguous numpy array; returns the bitmask."""
    assert data.flags["C_CONTIGUOUS"]
    n_rows, v = data.shape
    bitmask = np.zeros((n_rows + 31) // 32, dtype=np.int32)
    lib().oracle_compress(data.itemsize, _ptr(data), _ptr(data), n_rows, v, _ptr(_c(mask)), _ptr(_c(bitval)),
                          None, n_rows, _ptr(bitmask))
    return bitmask


def decompress_fetch(comp, mask, bitval, bitmask, index=None, clusters=None):
    """clusters: mask row of each stored row (mask/bitval [K x V])."""
    comp = _c(comp)
    n_rows, v = comp.shape
    idx = None if index is None else _c(np.asarray(index, dtype=np.int64))
    n = n_rows if idx is None else idx.shape[0]
    out = np.zeros((n, v), dtype=comp.dtype)
    lib().oracle_decompress_fetch_clustered(comp.itemsize, _ptr(comp), n_rows, v, _ptr(_c(mask)),
                                            _ptr(_c(bitval)), _ptr(_clusters(clusters)),
                                            _ptr(_c(np.asarray(bitmask, dtype=np.int32))), _ptr(idx), n, _ptr(out))
    return out


def compress_condensed(data, mask, bitval, index=None, clusters=None):
    """Returns (stream, offsets): stream of dtype data.dtype holding the rows
    back to back, offsets int64[n] the inclusive end byte offset of each row."""
    data = _c(data)
    n_rows, v = data.shape
    w = data.itemsize
    idx = None if index is None else _c(np.asarray(index, dtype=np.int64))
    n = n_rows if idx is None else idx.shape[0]
    offsets = np.zeros(n, dtype=np.int64)
    cl = _clusters(clusters)
    args = (_ptr(_c(mask)), _ptr(_c(bitval)), _ptr(cl), _ptr(idx), n)
    total = lib().oracle_compress_condensed(w, _ptr(data), n_rows, v, *args, None, _ptr(offsets))
    out = np.zeros(total // w, dtype=data.dtype)
    lib().oracle_compress_condensed(w, _ptr(data), n_rows, v, *args, _ptr(out), _ptr(offsets))
    return out, offsets


def decompress_condensed(stream, offsets, vec_size, mask, bitval, index=None, clusters=None):
    stream = _c(stream)
    offsets = _c(np.asarray(offsets, dtype=np.int64))
    idx = None if index is None else _c(np.asarray(index, dtype=np.int64))
    n = offsets.shape[0] if idx is None else idx.shape[0]
    out = np.zeros((n, vec_size), dtype=stream.dtype)
    lib().oracle_decompress_condensed(stream.itemsize, _ptr(stream), _ptr(offsets), offsets.shape[0], vec_size,
                                      _ptr(_c(mask)), _ptr(_c(bitval)), _ptr(_clusters(clusters)), _ptr(idx), n,
                                      _ptr(out))
    return out


def set_threads(n):
    """Threads of the oracle's OpenMP loops (count, score, k-means, compress_mt)."""
    lib().oracle_set_threads(int(n))


def random_shuffle(n, seed=None):
    a = np.arange(n, dtype=np.int32)
    if seed is not None:
        ctypes.CDLL(None).srand(ctypes.c_uint(seed))
    lib().oracle_random_shuffle(_ptr(a), n)
    return a


def kmeans(data, num_clusters, seed=1):
    data = _c(data)
    n, v = data.shape
    dt = np.dtype("u%d" % data.itemsize)
    masks = np.zeros((num_clusters, v), dtype=dt)
    bitvals = np.zeros((num_clusters, v), dtype=dt)
    clusters = np.zeros(n, dtype=np.int32)
    lib().oracle_kmeans(data.itemsize, _ptr(data), n, v, int(num_clusters), -1 if seed is None else int(seed),
                        _ptr(masks), _ptr(bitvals), _ptr(clusters))
    return masks, bitvals, clusters


def sweep_last():
    steps = ctypes.c_int(0)
    t = lib().oracle_sweep_last(ctypes.byref(steps))
    return t, steps.value
