"""Python host layer over the C ABI of libedpp.so (include/edpp.h).

The package directory name contains a hyphen, so import it with
``importlib.import_module("efficient-dpp_amd")``.  Device buffers are any
objects exposing ``data_ptr()`` (torch tensors on the GPU) or plain integer
addresses; torch is used only for allocation and streams.  There is no CPU
fallback: when libedpp.so is missing every call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libedpp.so")

# ---- status codes and enums (mirror include/edpp.h) -----------------------
OK, EINVAL, ENOMEM, ETIMEOUT, EHIP, ENODEV, EUNSUP = 0, -1, -2, -3, -4, -5, -6

SCAN_AUTO, SCAN_LOOKBACK, SCAN_RTS, SCAN_SSA, SCAN_PERSIST = 0, 1, 2, 3, 4
MATRIX_LM, MATRIX_REG, MATRIX_LM_REG, MATRIX_LM_SERIAL = 0, 1, 2, 3
NET_SERIAL, NET_BLELLOCH, NET_SKLANSKY, NET_WAVE = 0, 1, 2, 3
NET_SERIAL_MT, NET_BLELLOCH_CF, NET_GLOBAL_BLELLOCH, NET_GLOBAL_SKLANSKY, NET_GLOBAL_WAVE = 4, 5, 6, 7, 8
SPLIT_AUTO, SPLIT_WG, SPLIT_WI, SPLIT_SINGLE, SPLIT_CHAINED = 0, 1, 2, 3, 4
STREAM_COPY, STREAM_SCALE, STREAM_ADD, STREAM_TRIAD = 0, 1, 2, 3
PERM_RANDOM, PERM_BLOCK_LOCAL, PERM_IDENTITY = 0, 1, 2
KEYS_UNIFORM, KEYS_ZIPF = 0, 1
ACCESS_COLUMN, ACCESS_ROW, ACCESS_MIXED = 0, 1, 2

_lib = None

c_i32, c_i64, c_u64, c_int, c_vp, c_dp = (ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.POINTER(ctypes.c_double))

# name -> argtypes (all return int unless listed in _RESTYPES)
_SIGS = {
    "edpp_abi_version": [],
    "edpp_strerror": [c_int],
    "edpp_last_error": [c_vp],
    "edpp_ctx_create": [c_int, c_vp, ctypes.POINTER(c_vp)],
    "edpp_ctx_destroy": [c_vp],
    "edpp_ctx_sync": [c_vp],
    "edpp_ctx_set_watchdog": [c_vp, ctypes.c_uint32],
    "edpp_ctx_stream": [c_vp],
    "edpp_ctx_device_info": [c_vp, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_i64),
                             ctypes.POINTER(c_i64)],
    "edpp_stream_i32": [c_vp, c_int, c_vp, c_vp, c_vp, c_i64, c_dp],
    "edpp_gather_i32": [c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_dp],
    "edpp_scatter_i32": [c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_dp],
    "edpp_gather_i64": [c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_dp],
    "edpp_scatter_i64": [c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_dp],
    "edpp_default_passes": [c_vp, c_i64, c_int, c_int, ctypes.POINTER(c_int)],
    "edpp_hash_join_count": [c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_dp],
    "edpp_access_scale_i32": [c_vp, c_int, c_vp, c_vp, c_i64, c_int, c_dp],
    "edpp_access_chunks_i32": [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_dp],
    "edpp_access_random_i32": [c_vp, c_vp, c_i64, c_i64, c_u64, c_int, c_vp, c_dp],
    "edpp_sort_u32": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_dp],
    "edpp_sort_i32": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_dp],
    "edpp_scan_exclusive_i32": [c_vp, c_vp, c_vp, c_i64, c_int, c_vp, c_dp],
    "edpp_scan_exclusive_i64": [c_vp, c_vp, c_vp, c_i64, c_int, c_vp, c_dp],
    "edpp_scan_exclusive_i32_noalloc": [c_vp, c_vp, c_vp, c_i64, c_vp, c_dp],
    "edpp_scan_local_i32": [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_dp],
    "edpp_split_i32": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_dp],
    "edpp_split_i64": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_dp],
    "edpp_gen_i32": [c_vp, c_vp, c_i64, c_u64, c_i32],
    "edpp_gen_i64": [c_vp, c_vp, c_i64, c_u64, c_i64],
    "edpp_gen_perm": [c_vp, c_vp, c_i64, c_u64, c_int],
    "edpp_gen_split_keys_i32": [c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_u64],
    "edpp_gen_split_keys_i64": [c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_u64],
}


class Tuning(ctypes.Structure):
    """edpp_tuning of include/edpp_tune.h (0 / -1 fields mean the default)."""
    _fields_ = [("gs_unroll", c_i32), ("gs_random_nt", c_i32), ("scan_vpt", c_i32), ("split_ipt", c_i32),
                ("split_tile_mult", c_i32), ("scan_block", c_i32), ("split_block", c_i32),
                ("split_reorder", c_i32), ("sort_strategy", c_i32), ("reserved", c_i32 * 7)]


TUNING_FIELDS = ("gs_unroll", "gs_random_nt", "scan_vpt", "split_ipt", "split_tile_mult", "scan_block",
                 "split_block", "split_reorder", "sort_strategy")
REORDER_DEFAULT, REORDER_NONE, REORDER_FIXED, REORDER_VARIED = 0, 1, 2, 3

_SIGS["edpp_set_tuning"] = [c_vp, ctypes.POINTER(Tuning)]
_SIGS["edpp_get_tuning"] = [c_vp, ctypes.POINTER(Tuning)]

JOB_SCAN, JOB_SPLIT, JOB_GATHER = 0, 1, 2


class Job(ctypes.Structure):
    """edpp_job of include/edpp_batch.h."""
    _fields_ = [("kind", c_i32), ("elem_bytes", c_i32), ("n", c_i64), ("buckets", c_i32), ("reserved", c_i32),
                ("in_", c_vp), ("out", c_vp), ("idx", c_vp), ("vals_in", c_vp), ("vals_out", c_vp), ("aux", c_vp)]


_SIGS["edpp_batch_create"] = [c_vp, ctypes.POINTER(Job), c_i32, ctypes.POINTER(c_vp)]
_SIGS["edpp_batch_run"] = [c_vp, c_dp]
_SIGS["edpp_batch_run_graph"] = [c_vp, c_dp]
_SIGS["edpp_batch_destroy"] = [c_vp]
_SIGS["edpp_batch_set_occupancy"] = [c_vp, c_i32]
_SIGS["edpp_batch_info"] = [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]
_SIGS["edpp_batch_run_plain"] = [c_vp, ctypes.POINTER(Job), c_i32, c_dp]
_SIGS["edpp_batch_num_shapes"] = []
_SIGS["edpp_batch_shape_info"] = [c_i32] + [ctypes.POINTER(c_i32)] * 5
_SIGS["edpp_batch_set_shape"] = [c_vp, c_i32]
_SIGS["edpp_batch_set_order"] = [c_vp, c_i32]
BATCH_ORDER_PHASES, BATCH_ORDER_JOBS, BATCH_ORDER_XCD = 0, 1, 2
BATCH_ORDERS = ("phases", "jobs", "xcd")
_SIGS["edpp_batch_get_shape"] = [c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32), ctypes.POINTER(c_i32)]
_SIGS["edpp_batch_autotune"] = [c_vp, c_i32, ctypes.POINTER(ctypes.c_double)]
BATCH_OCCUPANCIES = (2, 4, 8)

OP_SCAN, OP_GATHER, OP_SCATTER, OP_SPLIT = 0, 1, 2, 3
NUM_FEATURES, HIDDEN, MAX_CANDIDATES = 12, 64, 15


class Call(ctypes.Structure):
    """edpp_call of include/edpp_learn.h."""
    _fields_ = [("op", c_i32), ("elem_bytes", c_i32), ("n", c_i64), ("buckets", c_i32), ("reserved", c_i32),
                ("in_", c_vp), ("out", c_vp), ("idx", c_vp), ("vals_in", c_vp), ("vals_out", c_vp)]


_c_fp = ctypes.POINTER(ctypes.c_float)
_SIGS["edpp_num_candidates"] = []
_SIGS["edpp_candidate_info"] = [c_i32, ctypes.POINTER(c_i32), ctypes.c_char_p, c_i32]
_SIGS["edpp_call_run"] = [c_vp, ctypes.POINTER(Call), c_i32, c_dp]
_SIGS["edpp_autotune"] = [c_vp, ctypes.POINTER(Call), c_i32, ctypes.POINTER(ctypes.c_double)]
_SIGS["edpp_call_features"] = [c_vp, ctypes.POINTER(Call), c_vp]
_SIGS["edpp_default_candidate"] = [c_vp, ctypes.POINTER(Call), ctypes.POINTER(c_i32)]
_SIGS["edpp_model_choice"] = [c_vp, ctypes.POINTER(Call), ctypes.POINTER(c_i32)]
_SIGS["edpp_mlp_num_params"] = []
_SIGS["edpp_mlp_create"] = [c_vp, c_u64, ctypes.POINTER(c_vp)]
_SIGS["edpp_mlp_destroy"] = [c_vp]
_SIGS["edpp_mlp_get_params"] = [c_vp, _c_fp, c_i64]
_SIGS["edpp_mlp_set_params"] = [c_vp, _c_fp, c_i64]
_SIGS["edpp_mlp_forward"] = [c_vp, c_vp, c_vp, c_i32, c_vp, c_vp]
_SIGS["edpp_mlp_loss_grad"] = [c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp]
_SIGS["edpp_mlp_train"] = [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, ctypes.c_float, _c_fp]
_SIGS["edpp_ctx_set_model"] = [c_vp, c_vp]
_SIGS["edpp_mlp_memo"] = [c_vp, c_i32]

_RESTYPES = {"edpp_strerror": ctypes.c_char_p, "edpp_last_error": ctypes.c_char_p, "edpp_ctx_stream": c_vp}


def declared_symbols():
    """Every entry point this binding declares (checked against the headers by the tests)."""
    return sorted(_SIGS)


def lib():
    """Load libedpp.so once; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run efficient-dpp_amd/build.py (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, c_int)
        _lib = L
    return _lib


class EdppError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"edpp status {status}: {msg}")
        self.status = status


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _buf(x, count, esize, what):
    """Device address of buffer `x` once it is known to hold `count` elements
    of `esize` bytes: a contiguous GPU tensor of that element width and at
    least that many elements.  A wrong buffer would hand a host or
    out-of-range address to a kernel, so every mismatch raises ValueError
    before the library is called.  None stays None (an optional buffer) and
    a plain integer is taken as a raw device address the caller vouches for.
    """
    if x is None or isinstance(x, int):
        return x
    if not getattr(x, "is_cuda", False):
        raise ValueError(f"{what}: not a GPU tensor (device {getattr(x, 'device', type(x).__name__)})")
    if not x.is_contiguous():
        raise ValueError(f"{what}: not contiguous")
    if esize is not None and x.element_size() != esize:
        raise ValueError(f"{what}: {x.element_size()}-byte elements where {esize}-byte ones are needed")
    if count is not None and x.numel() < count:
        raise ValueError(f"{what}: {x.numel()} elements where {count} are needed")
    return x.data_ptr()


def _count(n, what):
    n = int(n)
    if n < 0:
        raise ValueError(f"{what}: n={n} < 0")
    return n


def _numel(x):
    return x.numel() if hasattr(x, "numel") else None


class Context:
    """One library context: a HIP stream, a workspace and look-back state.

    ``stream`` is a raw hipStream_t address (for example
    ``torch.cuda.current_stream().cuda_stream``); None (0) is the default stream.
    """

    def __init__(self, device=0, stream=None):
        self._L = lib()
        h = c_vp()
        rc = self._L.edpp_ctx_create(int(device), stream, ctypes.byref(h))
        if rc != OK:
            raise EdppError(rc, self._L.edpp_strerror(rc).decode())
        self.h = h
        cus, xcds, l2, mall = c_int(), c_int(), c_i64(), c_i64()
        self._L.edpp_ctx_device_info(self.h, ctypes.byref(cus), ctypes.byref(xcds), ctypes.byref(l2),
                                     ctypes.byref(mall))
        self.cus, self.xcds, self.l2_bytes, self.mall_bytes = cus.value, xcds.value, l2.value, mall.value

    def close(self):
        if getattr(self, "h", None):
            self._L.edpp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- plumbing
    def _check(self, rc):
        if rc != OK:
            raise EdppError(rc, self._L.edpp_last_error(self.h).decode())

    def _call(self, name, *args, timed=False):
        ms = ctypes.c_double(0.0) if timed else None
        self._check(getattr(self._L, name)(self.h, *args, ctypes.byref(ms) if timed else None))
        return ms.value if timed else None

    def set_tuning(self, **kw):
        """Set schedule knobs (see include/edpp_tune.h); no arguments restores the defaults."""
        if not kw:
            self._check(self._L.edpp_set_tuning(self.h, None))
            return
        t = Tuning()
        t.gs_random_nt = -1
        for k, v in kw.items():
            if k not in TUNING_FIELDS:
                raise KeyError(k)
            setattr(t, k, int(v))
        self._check(self._L.edpp_set_tuning(self.h, ctypes.byref(t)))

    def get_tuning(self):
        t = Tuning()
        self._check(self._L.edpp_get_tuning(self.h, ctypes.byref(t)))
        return {k: getattr(t, k) for k in TUNING_FIELDS}

    # -- schedules, autotuning and the cost model (include/edpp_learn.h)
    def call_run(self, call, cand=-1, timed=False):
        ms = ctypes.c_double(0.0)
        self._check(self._L.edpp_call_run(self.h, ctypes.byref(call), int(cand), ctypes.byref(ms) if timed else None))
        return ms.value if timed else None

    def autotune(self, call, reps=3):
        """Median ms of every candidate of the call's family (-1 for other families)."""
        out = (ctypes.c_double * MAX_CANDIDATES)()
        self._check(self._L.edpp_autotune(self.h, ctypes.byref(call), int(reps), out))
        return list(out)

    def default_candidate(self, call):
        """Candidate index of the built-in schedule for this call (-1: none)."""
        k = c_i32(-1)
        self._check(self._L.edpp_default_candidate(self.h, ctypes.byref(call), ctypes.byref(k)))
        return k.value

    def model_choice(self, call):
        """Candidate index the context's model picks for this call."""
        k = c_i32(-1)
        self._check(self._L.edpp_model_choice(self.h, ctypes.byref(call), ctypes.byref(k)))
        return k.value

    def call_features(self, call, d_features):
        self._check(self._L.edpp_call_features(self.h, ctypes.byref(call),
                                               _buf(d_features, NUM_FEATURES, 4, "features")))

    def set_model(self, mlp):
        self._check(self._L.edpp_ctx_set_model(self.h, mlp.h if mlp is not None else None))

    def sync(self):
        self._check(self._L.edpp_ctx_sync(self.h))

    def set_watchdog(self, polls=1 << 22):
        """Polls of a device-side wait before it gives up and the next sync
        raises EdppError(ETIMEOUT); the default restores the library's."""
        self._check(self._L.edpp_ctx_set_watchdog(self.h, int(polls)))

    @property
    def stream(self):
        return self._L.edpp_ctx_stream(self.h)

    # -- primitives (n defaults to the element count of the first buffer).
    # Every tensor is checked by _buf before the call: on the GPU, contiguous,
    # of the element width the entry point reads, and long enough for n.
    def stream_op(self, op, a, b, c, n=None, timed=False):
        n = _count(_numel(a) if n is None else n, "stream")
        return self._call("edpp_stream_i32", int(op), _buf(a, n, 4, "stream a"), _buf(b, n, 4, "stream b"),
                          _buf(c, n, 4, "stream out"), n, timed=timed)

    def gather(self, src, dst, idx, n=None, passes=0, timed=False, wide=False):
        n = _count(_numel(idx) if n is None else n, "gather")
        name = "edpp_gather_i64" if wide else "edpp_gather_i32"
        w = 8 if wide else 4
        # src is read at positions idx[i] < its length, which the kernel checks
        # against n: it must hold n elements as well
        return self._call(name, _buf(src, n, w, "gather src"), _buf(dst, n, w, "gather dst"),
                          _buf(idx, n, 4, "gather idx"), n, int(passes), timed=timed)

    def scatter(self, src, dst, idx, n=None, passes=0, timed=False, wide=False):
        n = _count(_numel(idx) if n is None else n, "scatter")
        name = "edpp_scatter_i64" if wide else "edpp_scatter_i32"
        w = 8 if wide else 4
        return self._call(name, _buf(src, n, w, "scatter src"), _buf(dst, n, w, "scatter dst"),
                          _buf(idx, n, 4, "scatter idx"), n, int(passes), timed=timed)

    def sort(self, keys_in, keys_out, vals_in=None, vals_out=None, n=None, begin_bit=0, end_bit=32, signed=False,
             timed=False):
        """Stable LSD radix sort of 32-bit keys by bits [begin_bit, end_bit); values follow."""
        n = _count(_numel(keys_in) if n is None else n, "sort")
        name = "edpp_sort_i32" if signed else "edpp_sort_u32"
        return self._call(name, _buf(keys_in, n, 4, "sort keys_in"), _buf(keys_out, n, 4, "sort keys_out"),
                          _buf(vals_in, n, 4, "sort vals_in"), _buf(vals_out, n, 4, "sort vals_out"), n,
                          int(begin_bit), int(end_bit), timed=timed)

    def access_scale(self, pattern, src, dst, repeat, n=None, timed=False):
        """dst = 3 * src through the pattern's lane mapping (n a multiple of 64*repeat)."""
        n = _count(_numel(src) if n is None else n, "access_scale")
        return self._call("edpp_access_scale_i32", int(pattern), _buf(src, n, 4, "access src"),
                          _buf(dst, n, 4, "access dst"), n, int(repeat), timed=timed)

    def access_chunks(self, src, chunk, order, count, n=None, timed=False):
        """count[0] = odd elements of src, read in chunks taken in `order`."""
        n = _count(_numel(src) if n is None else n, "access_chunks")
        chunks = -(-n // int(chunk)) if int(chunk) > 0 else None
        return self._call("edpp_access_chunks_i32", _buf(src, n, 4, "access src"), n, int(chunk),
                          _buf(order, chunks, 4, "access order"), _buf(count, 1, 8, "access count"), timed=timed)

    def access_random(self, table, accesses, seed, write=False, total=None, table_n=None, timed=False):
        """Random-access probe: `accesses` reads (total[0] = their sum) or
        writes of table[fmix32(i + seed) & (table_n - 1)]."""
        table_n = _count(_numel(table) if table_n is None else table_n, "access_random")
        return self._call("edpp_access_random_i32", _buf(table, table_n, 4, "access table"), table_n, int(accesses),
                          seed & (2**64 - 1), int(bool(write)), _buf(total, 1, 8, "access total"), timed=timed)

    def join_count(self, r_keys, s_keys, count, r_n=None, s_n=None, timed=False):
        """count[0] (uint64/int64 device tensor) = number of equal-key pairs of R and S."""
        r_n = _count(_numel(r_keys) if r_n is None else r_n, "join R")
        s_n = _count(_numel(s_keys) if s_n is None else s_n, "join S")
        return self._call("edpp_hash_join_count", _buf(r_keys, r_n, 4, "join R"), r_n, _buf(s_keys, s_n, 4, "join S"),
                          s_n, _buf(count, 1, 8, "join count"), timed=timed)

    def schedule_passes(self, n, scatter=False, wide=False):
        """Pass count a passes=0 gather/scatter of n elements uses without a cost model."""
        out = c_int(0)
        self._check(self._L.edpp_default_passes(self.h, int(n), 8 if wide else 4, int(bool(scatter)),
                                                ctypes.byref(out)))
        return out.value

    def scan(self, src, dst, n=None, scheme=SCAN_AUTO, total=None, timed=False, wide=False):
        n = _count(_numel(src) if n is None else n, "scan")
        name = "edpp_scan_exclusive_i64" if wide else "edpp_scan_exclusive_i32"
        w = 8 if wide else 4
        return self._call(name, _buf(src, n, w, "scan in"), _buf(dst, n, w, "scan out"), n, int(scheme),
                          _buf(total, 1, w, "scan total"), timed=timed)

    def scan_noalloc(self, src, dst, n=None, total=None, timed=False):
        """Exclusive int32 scan that allocates nothing: its look-back state is
        the library's fixed per-GPU scratch array (calls on one GPU must not
        overlap)."""
        n = _count(_numel(src) if n is None else n, "scan_noalloc")
        return self._call("edpp_scan_exclusive_i32_noalloc", _buf(src, n, 4, "scan in"), _buf(dst, n, 4, "scan out"),
                          n, _buf(total, 1, 4, "scan total"), timed=timed)

    def scan_local(self, src, dst, tile_size, local_size=512, matrix=MATRIX_LM_REG, network=NET_WAVE, timed=False):
        n = max(int(tile_size), 0) * max(int(local_size), 0)
        return self._call("edpp_scan_local_i32", _buf(src, n, 4, "scan_local in"), _buf(dst, n, 4, "scan_local out"),
                          int(tile_size), int(local_size), int(matrix), int(network), timed=timed)

    def split(self, keys_in, keys_out, buckets, vals_in=None, vals_out=None, start=None, n=None,
              strategy=SPLIT_AUTO, timed=False, wide=False):
        n = _count(_numel(keys_in) if n is None else n, "split")
        name = "edpp_split_i64" if wide else "edpp_split_i32"
        w = 8 if wide else 4
        return self._call(name, _buf(keys_in, n, w, "split keys_in"), _buf(keys_out, n, w, "split keys_out"),
                          _buf(vals_in, n, 4, "split vals_in"), _buf(vals_out, n, 4, "split vals_out"),
                          _buf(start, max(int(buckets), 0), 4, "split start"), n, int(buckets), int(strategy),
                          timed=timed)

    # -- reproducible inputs
    def gen_i32(self, out, seed, mask=-1, n=None):
        n = _count(_numel(out) if n is None else n, "gen_i32")
        self._check(self._L.edpp_gen_i32(self.h, _buf(out, n, 4, "gen out"), n, seed & (2**64 - 1), mask))

    def gen_i64(self, out, seed, mask=-1, n=None):
        n = _count(_numel(out) if n is None else n, "gen_i64")
        self._check(self._L.edpp_gen_i64(self.h, _buf(out, n, 8, "gen out"), n, seed & (2**64 - 1), mask))

    def gen_perm(self, idx, seed, kind=PERM_RANDOM, n=None):
        n = _count(_numel(idx) if n is None else n, "gen_perm")
        self._check(self._L.edpp_gen_perm(self.h, _buf(idx, n, 4, "gen_perm idx"), n, seed & (2**64 - 1), int(kind)))

    def gen_split_keys(self, keys, vals, buckets, dist, seed, n=None, wide=False):
        n = _count(_numel(keys) if n is None else n, "gen_split_keys")
        name = "edpp_gen_split_keys_i64" if wide else "edpp_gen_split_keys_i32"
        self._check(getattr(self._L, name)(self.h, _buf(keys, n, 8 if wide else 4, "gen keys"),
                                           _buf(vals, n, 4, "gen vals"), n, int(buckets), int(dist),
                                           seed & (2**64 - 1)))


def make_job(kind, n, src, dst, elem_bytes=4, idx=None, buckets=0, vals_in=None, vals_out=None, aux=None):
    """One batch job descriptor; buffers as for the single-call methods and
    checked the same way (aux: the scan total, or the split's bucket starts)."""
    n = _count(n, "job")
    w = int(elem_bytes) if int(elem_bytes) in (4, 8) else None
    aux_count, aux_w = (1, w) if int(kind) == JOB_SCAN else ((max(int(buckets), 0), 4) if int(kind) == JOB_SPLIT
                                                             else (None, None))
    return Job(int(kind), int(elem_bytes), n, int(buckets), 0, _buf(src, n, w, "job in"), _buf(dst, n, w, "job out"),
               _buf(idx, n, 4, "job idx"), _buf(vals_in, n, 4, "job vals_in"), _buf(vals_out, n, 4, "job vals_out"),
               _buf(aux, aux_count, aux_w, "job aux"))


class Batch:
    """A planned batch: one persistent launch per run (include/edpp_batch.h)."""

    def __init__(self, ctx, jobs):
        self.ctx = ctx
        self._jobs = (Job * len(jobs))(*jobs)
        h = c_vp()
        ctx._check(ctx._L.edpp_batch_create(ctx.h, self._jobs, len(jobs), ctypes.byref(h)))
        self.h = h

    def run(self, timed=False, graph=False):
        ms = ctypes.c_double(0.0)
        f = self.ctx._L.edpp_batch_run_graph if graph else self.ctx._L.edpp_batch_run
        self.ctx._check(f(self.h, ctypes.byref(ms) if timed else None))
        return ms.value if timed else None

    def set_occupancy(self, wgs_per_cu):
        self.ctx._check(self.ctx._L.edpp_batch_set_occupancy(self.h, int(wgs_per_cu)))

    def set_shape(self, shape):
        self.ctx._check(self.ctx._L.edpp_batch_set_shape(self.h, int(shape)))

    def set_order(self, order):
        self.ctx._check(self.ctx._L.edpp_batch_set_order(self.h, int(order)))

    def get_shape(self):
        """(shape index, workgroups per compute unit, queue order) currently set."""
        s, o, q = c_i32(), c_i32(), c_i32()
        self.ctx._check(self.ctx._L.edpp_batch_get_shape(self.h, ctypes.byref(s), ctypes.byref(o), ctypes.byref(q)))
        return s.value, o.value, q.value

    def autotune(self, reps=3):
        """Times every queue order x shape x occupancy (median of reps); leaves
        the fastest set and returns {(shape, wgs_per_cu, order): ms}."""
        k = self.ctx._L.edpp_batch_num_shapes()
        no = len(BATCH_OCCUPANCIES)
        nq = len(BATCH_ORDERS)
        out = (ctypes.c_double * (nq * k * no))()
        self.ctx._check(self.ctx._L.edpp_batch_autotune(self.h, int(reps), out))
        return {(s, o, q): out[(q * k + s) * no + i] for q in range(nq) for s in range(k)
                for i, o in enumerate(BATCH_OCCUPANCIES)}

    def info(self):
        a, b, c = c_i64(), c_i64(), c_i64()
        self.ctx._check(self.ctx._L.edpp_batch_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"tiles": a.value, "hist_tiles": b.value, "hscan_tiles": c.value}

    def run_plain(self, timed=False):
        ms = ctypes.c_double(0.0)
        self.ctx._check(self.ctx._L.edpp_batch_run_plain(self.ctx.h, self._jobs, len(self._jobs),
                                                         ctypes.byref(ms) if timed else None))
        return ms.value if timed else None

    def close(self):
        if getattr(self, "h", None):
            self.ctx._L.edpp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_shapes():
    """[{block, scan_tile32, scan_tile64, split_tile, gather_tile}] of every batch tile shape."""
    L = lib()
    out = []
    for k in range(L.edpp_batch_num_shapes()):
        v = [c_i32() for _ in range(5)]
        L.edpp_batch_shape_info(k, *[ctypes.byref(x) for x in v])
        out.append(dict(zip(("block", "scan_tile32", "scan_tile64", "split_tile", "gather_tile"),
                            (x.value for x in v))))
    return out


def make_call(op, n, src, dst, elem_bytes=4, idx=None, buckets=0, vals_in=None, vals_out=None):
    """One call descriptor for the autotuner and the cost model, its buffers
    checked as for the single-call methods."""
    n = _count(n, "call")
    w = int(elem_bytes) if int(elem_bytes) in (4, 8) else None
    return Call(int(op), int(elem_bytes), n, int(buckets), 0, _buf(src, n, w, "call in"), _buf(dst, n, w, "call out"),
                _buf(idx, n, 4, "call idx"), _buf(vals_in, n, 4, "call vals_in"), _buf(vals_out, n, 4, "call vals_out"))


def candidates():
    """[(family, name)] of every schedule candidate."""
    L = lib()
    out = []
    for k in range(L.edpp_num_candidates()):
        fam = c_i32()
        name = ctypes.create_string_buffer(96)
        L.edpp_candidate_info(k, ctypes.byref(fam), name, 96)
        out.append((fam.value, name.value.decode()))
    return out


class MLP:
    """The FP32 cost model 12 -> 64 -> 64 -> 15 trained by the library's own HIP kernels."""

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        h = c_vp()
        ctx._check(ctx._L.edpp_mlp_create(ctx.h, seed & (2**64 - 1), ctypes.byref(h)))
        self.h = h
        self.nparams = ctx._L.edpp_mlp_num_params()

    def get_params(self):
        import numpy as np
        out = np.empty(self.nparams, np.float32)
        self.ctx._check(self.ctx._L.edpp_mlp_get_params(self.h, out.ctypes.data_as(_c_fp), self.nparams))
        return out

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, np.float32)
        self.ctx._check(self.ctx._L.edpp_mlp_set_params(self.h, p.ctypes.data_as(_c_fp), p.size))

    def forward(self, x, mask, probs=None, choice=None):
        rows = x.shape[0]
        self.ctx._check(self.ctx._L.edpp_mlp_forward(
            self.h, _buf(x, rows * NUM_FEATURES, 4, "mlp x"), _buf(mask, rows * MAX_CANDIDATES, 4, "mlp mask"), rows,
            _buf(probs, rows * MAX_CANDIDATES, 4, "mlp probs"), _buf(choice, rows, 4, "mlp choice")))

    def loss_grad(self, x, target, mask, loss, grad):
        rows = x.shape[0]
        self.ctx._check(self.ctx._L.edpp_mlp_loss_grad(
            self.h, _buf(x, rows * NUM_FEATURES, 4, "mlp x"), _buf(target, rows * MAX_CANDIDATES, 4, "mlp target"),
            _buf(mask, rows * MAX_CANDIDATES, 4, "mlp mask"), rows, _buf(loss, 1, 4, "mlp loss"),
            _buf(grad, self.nparams, 4, "mlp grad")))

    def train(self, x, target, mask, epochs, lr=1e-2):
        import numpy as np
        losses = np.zeros(max(epochs, 1), np.float32)
        rows = x.shape[0]
        self.ctx._check(self.ctx._L.edpp_mlp_train(
            self.h, _buf(x, rows * NUM_FEATURES, 4, "mlp x"), _buf(target, rows * MAX_CANDIDATES, 4, "mlp target"),
            _buf(mask, rows * MAX_CANDIDATES, 4, "mlp mask"), rows, int(epochs), float(lr),
            losses.ctypes.data_as(_c_fp)))
        return losses[:epochs]

    def memo(self, enabled=True):
        """Remember the model's choice per AUTO call signature (on by default);
        either setting forgets what was remembered."""
        self.ctx._check(self.ctx._L.edpp_mlp_memo(self.h, 1 if enabled else 0))

    def close(self):
        if getattr(self, "h", None):
            self.ctx._L.edpp_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
