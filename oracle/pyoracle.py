"""numpy front end of the C oracle (oracle.c) and of the OpenMP baselines.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64

_oracle = None
_baseline = None
_ref = None


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run oracle/build.py")
    return ctypes.CDLL(path)


def oracle():
    global _oracle
    if _oracle is None:
        L = _load("liboracle.so")
        sig = {
            "oracle_splitmix_at": (_U64, [_U64, _U64]),
            "oracle_splitmix_seq": (None, [_U64, _P, _I64]),
            "oracle_gen_i32": (None, [_P, _I64, _U64, ctypes.c_int32]),
            "oracle_gen_i64": (None, [_P, _I64, _U64, ctypes.c_int64]),
            "oracle_gen_perm": (None, [_P, _I64, _U64, ctypes.c_int]),
            "oracle_zipf_thresholds": (None, [ctypes.c_int, _P]),
            "oracle_gen_split_keys_i32": (None, [_P, _P, _I64, ctypes.c_int, ctypes.c_int, _U64]),
            "oracle_gen_split_keys_i64": (None, [_P, _P, _I64, ctypes.c_int, ctypes.c_int, _U64]),
            "oracle_scan_exclusive_u32": (ctypes.c_uint32, [_P, _P, _I64]),
            "oracle_scan_exclusive_u64": (ctypes.c_uint64, [_P, _P, _I64]),
            "oracle_gather": (None, [_P, _P, _P, _I64, ctypes.c_int, ctypes.c_int]),
            "oracle_scatter": (None, [_P, _P, _P, _I64, ctypes.c_int, ctypes.c_int]),
            "oracle_split": (None, [_P, _P, _P, _P, _P, _I64, ctypes.c_int, ctypes.c_int]),
            "oracle_sort_u32": (None, [_P, _P, _P, _P, _I64, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "oracle_join_count": (_U64, [_P, _I64, _P, _I64]),
            "oracle_stream": (None, [ctypes.c_int, _P, _P, _P, _I64]),
            "oracle_hampel_mean": (ctypes.c_double, [_P, ctypes.c_int]),
        }
        for k, (res, args) in sig.items():
            f = getattr(L, k)
            f.restype = res
            f.argtypes = args
        _oracle = L
    return _oracle


def baseline():
    global _baseline
    if _baseline is None:
        L = _load("libbaseline.so")
        for k in ["baseline_scan_ssa", "baseline_scan_rts", "baseline_copy", "baseline_scale"]:
            getattr(L, k).restype = ctypes.c_double
            getattr(L, k).argtypes = [_P, _P, _I64]
        for k in ["baseline_gather", "baseline_scatter"]:
            getattr(L, k).restype = ctypes.c_double
            getattr(L, k).argtypes = [_P, _P, _P, _I64]
        L.baseline_threads.restype = ctypes.c_int
        L.baseline_fill.argtypes = [_P, _I64, ctypes.c_int32]
        _baseline = L
    return _baseline


def reference():
    """The reference's own OpenMP functions (oracle/_ref), or None if not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libedpp_ref_openmp.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        for k in ["edpp_ref_gather", "edpp_ref_scatter"]:
            getattr(L, k).restype = ctypes.c_double
            getattr(L, k).argtypes = [_P, _P, _P, _U64]
        for k in ["edpp_ref_copy", "edpp_ref_scale", "edpp_ref_copy_ss", "edpp_ref_scale_ss"]:
            getattr(L, k).restype = ctypes.c_double
            getattr(L, k).argtypes = [_P, _P, _U64]
        L.edpp_ref_compute_bandwidth.restype = ctypes.c_double
        L.edpp_ref_compute_bandwidth.argtypes = [_U64, ctypes.c_int, ctypes.c_double]
        L.edpp_ref_average_hampel.restype = ctypes.c_double
        L.edpp_ref_average_hampel.argtypes = [_P, ctypes.c_int]
        _ref = L
    return _ref


def _a(x):
    return x.ctypes.data_as(_P)


# ---- generators ------------------------------------------------------------
def splitmix_seq(seed, n):
    out = np.empty(n, np.uint64)
    oracle().oracle_splitmix_seq(seed & (2**64 - 1), _a(out), n)
    return out


def gen_i32(n, seed, mask=-1):
    out = np.empty(n, np.int32)
    oracle().oracle_gen_i32(_a(out), n, seed & (2**64 - 1), mask)
    return out


def gen_i64(n, seed, mask=-1):
    out = np.empty(n, np.int64)
    oracle().oracle_gen_i64(_a(out), n, seed & (2**64 - 1), mask)
    return out


def gen_perm(n, seed, kind=0):
    out = np.empty(n, np.int32)
    oracle().oracle_gen_perm(_a(out), n, seed & (2**64 - 1), kind)
    return out


def zipf_thresholds(B):
    out = np.empty(B, np.uint64)
    oracle().oracle_zipf_thresholds(B, _a(out))
    return out


def gen_split_keys(n, B, dist, seed, wide=False):
    keys = np.empty(n, np.int64 if wide else np.int32)
    vals = np.empty(n, np.int32)
    f = oracle().oracle_gen_split_keys_i64 if wide else oracle().oracle_gen_split_keys_i32
    f(_a(keys), _a(vals), n, B, dist, seed & (2**64 - 1))
    return keys, vals


# ---- primitives --------------------------------------------------------------
def scan_exclusive(x):
    """Exclusive scan with wrap-around; returns (out, total) in x's dtype."""
    x = np.ascontiguousarray(x)
    if x.dtype.itemsize == 4:
        u = x.view(np.uint32)
        out = np.empty_like(u)
        tot = oracle().oracle_scan_exclusive_u32(_a(u), _a(out), u.size)
    else:
        u = x.view(np.uint64)
        out = np.empty_like(u)
        tot = oracle().oracle_scan_exclusive_u64(_a(u), _a(out), u.size)
    return out.view(x.dtype), np.array(tot, dtype=u.dtype).view(x.dtype)[()]


def gather(src, idx, passes=1, out=None):
    src = np.ascontiguousarray(src)
    out = np.zeros(idx.size, src.dtype) if out is None else out
    oracle().oracle_gather(_a(src), _a(out), _a(np.ascontiguousarray(idx, np.int32)), idx.size, passes,
                           src.dtype.itemsize)
    return out


def scatter(src, idx, passes=1, out=None):
    src = np.ascontiguousarray(src)
    out = np.zeros(idx.size, src.dtype) if out is None else out
    oracle().oracle_scatter(_a(src), _a(out), _a(np.ascontiguousarray(idx, np.int32)), idx.size, passes,
                            src.dtype.itemsize)
    return out


def split(keys, B, vals=None):
    """Stable split; returns (keys_out, vals_out or None, start[B])."""
    keys = np.ascontiguousarray(keys)
    ko = np.empty_like(keys)
    vo = None if vals is None else np.empty_like(vals)
    start = np.empty(B, np.int32)
    oracle().oracle_split(_a(keys), _a(ko), None if vals is None else _a(np.ascontiguousarray(vals)),
                          None if vo is None else _a(vo), _a(start), keys.size, B, keys.dtype.itemsize)
    return ko, vo, start


def sort(keys, vals=None, begin_bit=0, end_bit=32, signed=False):
    """Stable LSD radix sort of 32-bit keys by bits [begin_bit, end_bit);
    returns (keys_out, vals_out or None)."""
    keys = np.ascontiguousarray(keys)
    assert keys.dtype.itemsize == 4
    ko = np.empty_like(keys)
    vo = None if vals is None else np.empty_like(vals)
    oracle().oracle_sort_u32(_a(keys), None if vals is None else _a(np.ascontiguousarray(vals)), _a(ko),
                             None if vo is None else _a(vo), keys.size, begin_bit, end_bit, int(signed))
    return ko, vo


def join_count(r, s):
    """Number of pairs (i, j) with r[i] == s[j] (int32 keys)."""
    r = np.ascontiguousarray(r, np.int32)
    s = np.ascontiguousarray(s, np.int32)
    return int(oracle().oracle_join_count(_a(r), r.size, _a(s), s.size))


def stream(op, a, b=None):
    c = np.empty_like(a)
    oracle().oracle_stream(op, _a(a), None if b is None else _a(b), _a(c), a.size)
    return c


def hampel_mean(x):
    x = np.ascontiguousarray(x, np.float64)
    return oracle().oracle_hampel_mean(_a(x), x.size)
