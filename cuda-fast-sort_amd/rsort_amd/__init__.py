"""Python host mirror of rsort's API over the C ABI of librsort_amd.so.

Same four entry points and argument meaning as the reference templates
(include/rsort/rsort.cuh:13-92):

    temp = sort(d_inout, None, None, n)           # size query (no device work)
    sort(d_inout, temp, d_workspace, n)           # in-place ascending sort
    sort_descending(...), sort_pairs(keys, values, temp, ws, n), sort_pairs_descending(...)

Device buffers are torch tensors on the GPU (torch provides memory and
streams only).  The library is loaded from this source tree; if it is
missing the import fails loudly -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# RSORT_AMD_LIB selects another in-tree build (e.g. librsort_amd_timing.so).
LIB_PATH = os.path.join(PKG_DIR, os.environ.get("RSORT_AMD_LIB", "librsort_amd.so"))

KEY_UNSIGNED, KEY_SIGNED, KEY_FLOAT = 0, 1, 2
FLAG_NO_PASS_SKIP = 1
FLAG_GRAPH = 2  # replay identical calls from a cached hipGraph

SUCCESS = 0
STATUS_NAMES = {
    0: "RSORT_SUCCESS",
    1: "RSORT_ERROR_INVALID_ARGUMENT",
    2: "RSORT_ERROR_UNSUPPORTED",
    3: "RSORT_ERROR_WORKSPACE_TOO_SMALL",
    4: "RSORT_ERROR_DEVICE",
    5: "RSORT_ERROR_BUSY",
}


class RsortError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Desc(ctypes.Structure):
    _fields_ = [
        ("key_kind", ctypes.c_uint32),
        ("key_bytes", ctypes.c_uint32),
        ("value_bytes", ctypes.c_uint32),
        ("descending", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("passes", ctypes.c_uint32),
        ("threads", ctypes.c_uint32),
        ("items_per_thread", ctypes.c_uint32),
        ("tile", ctypes.c_uint32),
        ("num_tiles", ctypes.c_uint64),
        ("wide", ctypes.c_uint32),
        ("index_mode", ctypes.c_uint32),
        ("workspace_bytes", ctypes.c_uint64),
        ("lookback_bytes", ctypes.c_uint64),
    ]


SEGMENT_BINS = 9  # RSORT_SEGMENT_BINS


class SegPlanInfo(ctypes.Structure):
    _fields_ = [
        ("bin_capacity", ctypes.c_uint32 * SEGMENT_BINS),
        ("tile", ctypes.c_uint32),
        ("max_long_segments", ctypes.c_uint64),
        ("max_tiles", ctypes.c_uint64),
        ("wide", ctypes.c_uint32),
        ("index_mode", ctypes.c_uint32),
        ("workspace_bytes", ctypes.c_uint64),
    ]


class ServeConfig(ctypes.Structure):
    _fields_ = [
        ("desc", Desc),
        ("batch_keys", ctypes.c_uint64),
        ("batch_requests", ctypes.c_uint32),
        ("window_us", ctypes.c_uint32),
        ("depth", ctypes.c_uint32),
        ("max_pending", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [("latency_us", ctypes.c_uint64), ("queue_us", ctypes.c_uint64), ("batch", ctypes.c_uint64),
                ("n", ctypes.c_uint64)]


class ServeStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("completed_keys", ctypes.c_uint64),
                ("pending", ctypes.c_uint64), ("in_flight", ctypes.c_uint64), ("graphs", ctypes.c_uint32),
                ("direct", ctypes.c_uint32), ("workspace_bytes", ctypes.c_uint64)]


class Tuning(ctypes.Structure):
    _fields_ = [("tile_candidate", ctypes.c_uint32), ("hist_blocks_per_cu", ctypes.c_uint32),
                ("lds_threshold", ctypes.c_uint32), ("small_tile_candidate", ctypes.c_uint32),
                ("small_log2n", ctypes.c_uint32), ("tiny_tile_candidate", ctypes.c_uint32),
                ("tiny_log2n", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


SERVE_NO_GRAPH = 1
SERVE_BATCHED = 2

_lib = None


def lib() -> ctypes.CDLL:
    """Load librsort_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with python cuda-fast-sort_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.rsort_sort.argtypes = [ctypes.POINTER(Desc), vp, vp, vp, ctypes.POINTER(ctypes.c_size_t), u64, vp]
        L.rsort_sort.restype = i32
        L.rsort_sort_noalloc.argtypes = [ctypes.POINTER(Desc), vp, vp, vp, vp, u64, vp]
        L.rsort_sort_noalloc.restype = i32
        L.rsort_stream_bandwidth.argtypes = [ctypes.c_uint32, vp, vp, u64, ctypes.c_uint32, vp]
        L.rsort_stream_bandwidth.restype = i32
        L.rsort_sort_profiled.argtypes = [ctypes.POINTER(Desc), vp, vp, vp, ctypes.POINTER(ctypes.c_size_t), u64, vp,
                                          ctypes.POINTER(ctypes.c_float), i32, ctypes.POINTER(i32)]
        L.rsort_sort_profiled.restype = i32
        L.rsort_plan.argtypes = [ctypes.POINTER(Desc), u64, ctypes.POINTER(PlanInfo)]
        L.rsort_plan.restype = i32
        u32 = ctypes.c_uint32
        L.rsort_sort_records.argtypes = [ctypes.POINTER(Desc), vp, u64, u32, u32, vp, ctypes.POINTER(ctypes.c_size_t),
                                         vp]
        L.rsort_sort_records.restype = i32
        L.rsort_segmented_sort.argtypes = [ctypes.POINTER(Desc), vp, vp, u64, vp, u32, u64, vp,
                                           ctypes.POINTER(ctypes.c_size_t), vp]
        L.rsort_segmented_sort.restype = i32
        L.rsort_segmented_plan.argtypes = [ctypes.POINTER(Desc), u64, u32, u64, ctypes.POINTER(SegPlanInfo)]
        L.rsort_segmented_plan.restype = i32
        L.rsort_serve_create.argtypes = [ctypes.POINTER(ServeConfig), ctypes.POINTER(vp)]
        L.rsort_serve_submit.argtypes = [vp, vp, vp, u64, ctypes.POINTER(ctypes.c_uint64)]
        L.rsort_serve_wait.argtypes = [vp, u64, u64, ctypes.POINTER(RequestInfo)]
        L.rsort_serve_wait_many.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), u64, u64, ctypes.POINTER(RequestInfo)]
        L.rsort_serve_submit_many.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                              ctypes.POINTER(ctypes.c_uint64), u64, ctypes.POINTER(ctypes.c_uint64)]
        L.rsort_serve_flush.argtypes = [vp]
        L.rsort_serve_stats.argtypes = [vp, ctypes.POINTER(ServeStats)]
        L.rsort_serve_destroy.argtypes = [vp]
        L.rsort_init_keys.argtypes = [u32, u32, u32, vp, u64, vp]
        L.rsort_init_keys.restype = i32
        L.rsort_set_seed_radix.argtypes = [ctypes.c_int32]
        L.rsort_set_seed_radix.restype = u32
        L.rsort_check_sorted.argtypes = [ctypes.POINTER(Desc), vp, u64, vp, vp]
        L.rsort_check_sorted.restype = i32
        L.rsort_check_pairings.argtypes = [u32, u32, vp, vp, vp, u64, vp, vp]
        L.rsort_check_pairings.restype = i32
        L.rsort_digit_counts.argtypes = [ctypes.POINTER(Desc), vp, u64, vp, vp]
        L.rsort_digit_counts.restype = i32
        L.rsort_tuning_get.argtypes = [u32, u32, ctypes.POINTER(Tuning)]
        L.rsort_tuning_get.restype = i32
        L.rsort_tuning_set.argtypes = [u32, u32, ctypes.POINTER(Tuning)]
        L.rsort_tuning_set.restype = i32
        L.rsort_tuning_load.argtypes = [ctypes.c_char_p]
        L.rsort_tuning_load.restype = i32
        for f in ("create", "submit", "submit_many", "wait", "flush", "stats", "destroy"):
            getattr(L, "rsort_serve_" + f).restype = i32
        L.rsort_status_string.argtypes = [i32]
        L.rsort_status_string.restype = ctypes.c_char_p
        L.rsort_last_error.argtypes = []
        L.rsort_last_error.restype = ctypes.c_char_p
        L.rsort_abi_version.argtypes = []
        L.rsort_abi_version.restype = i32
        L.rsort_debug_phase_slots.argtypes = [ctypes.POINTER(ctypes.c_ulonglong), i32, i32]
        L.rsort_debug_phase_slots.restype = i32
        _lib = L
    return _lib


def last_error() -> str:
    return lib().rsort_last_error().decode()


def _check(status: int) -> None:
    if status != SUCCESS:
        raise RsortError(status, last_error())


@dataclass(frozen=True)
class KeyType:
    kind: int
    nbytes: int


def key_type_of(dtype) -> KeyType:
    """Key kind/width of a torch dtype (radix_traits, tuning.cuh:297-397)."""
    import torch

    table = {
        torch.uint8: KeyType(KEY_UNSIGNED, 1), torch.int8: KeyType(KEY_SIGNED, 1),
        torch.uint16: KeyType(KEY_UNSIGNED, 2), torch.int16: KeyType(KEY_SIGNED, 2),
        torch.uint32: KeyType(KEY_UNSIGNED, 4), torch.int32: KeyType(KEY_SIGNED, 4),
        torch.uint64: KeyType(KEY_UNSIGNED, 8), torch.int64: KeyType(KEY_SIGNED, 8),
        torch.float16: KeyType(KEY_FLOAT, 2), torch.bfloat16: KeyType(KEY_FLOAT, 2),
        torch.float32: KeyType(KEY_FLOAT, 4), torch.float64: KeyType(KEY_FLOAT, 8),
    }
    if dtype not in table:
        raise TypeError(f"no radix key mapping for {dtype}; pass key_type=KeyType(kind, nbytes)")
    return table[dtype]


def make_desc(key_kind: int, key_bytes: int, value_bytes: int = 0, descending: bool = False, flags: int = 0) -> Desc:
    return Desc(key_kind, key_bytes, value_bytes, 1 if descending else 0, flags, 0)


def _stream_ptr(stream) -> int | None:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


def sort_raw(desc: Desc, keys_ptr: int | None, vals_ptr: int | None, ws_ptr: int | None, temp_bytes: int | None,
             n: int, stream=None) -> int:
    """Direct call of rsort_sort; returns the workspace size (query) or temp_bytes."""
    tb = ctypes.c_size_t(temp_bytes or 0)
    _check(lib().rsort_sort(ctypes.byref(desc), keys_ptr, vals_ptr, ws_ptr, ctypes.byref(tb), n, _stream_ptr(stream)))
    return int(tb.value)


def sort_noalloc_raw(desc: Desc, keys_ptr: int | None, vals_ptr: int | None, keys_alt_ptr: int | None,
                     vals_alt_ptr: int | None, n: int, stream=None) -> None:
    """Direct call of rsort_sort_noalloc: sorts in place with the caller's
    alternate buffers as ping-pong partners and the library's per-device sort
    state; allocates nothing."""
    _check(lib().rsort_sort_noalloc(ctypes.byref(desc), keys_ptr, vals_ptr, keys_alt_ptr, vals_alt_ptr, n,
                                    _stream_ptr(stream)))


def stream_bandwidth(mode: int, dst_ptr: int | None, src_ptr: int | None, nbytes: int, blocks_per_cu: int = 8,
                     stream=None) -> None:
    """rsort_stream_bandwidth: one streaming launch (0 read, 1 write, 2 copy, 3 scatter copy)."""
    _check(lib().rsort_stream_bandwidth(mode, dst_ptr, src_ptr, nbytes, blocks_per_cu, _stream_ptr(stream)))


def sort_records_raw(desc: Desc, records_ptr: int | None, n: int, record_bytes: int, key_offset: int,
                     ws_ptr: int | None, temp_bytes: int | None, stream=None) -> int:
    """Direct call of rsort_sort_records; returns the workspace size (query) or temp_bytes."""
    tb = ctypes.c_size_t(temp_bytes or 0)
    _check(lib().rsort_sort_records(ctypes.byref(desc), records_ptr, n, record_bytes, key_offset, ws_ptr,
                                    ctypes.byref(tb), _stream_ptr(stream)))
    return int(tb.value)


def sort_records(records, record_bytes: int, key_offset: int, key_type: "KeyType", temp_bytes, d_workspace,
                 descending: bool = False, stream=None):
    """Array-of-structs sort: `records` is a device tensor holding n records of
    record_bytes bytes; sorts them stably in place by the key_type field at
    key_offset.  d_workspace None -> returns the workspace size."""
    nbytes = records.numel() * records.element_size()
    if nbytes % record_bytes:
        raise ValueError("records tensor is not a whole number of records")
    n = nbytes // record_bytes
    d = make_desc(key_type.kind, key_type.nbytes, 0, descending)
    if d_workspace is None:
        return sort_records_raw(d, None, n, record_bytes, key_offset, None, 0)
    return sort_records_raw(d, _ptr(records), n, record_bytes, key_offset, _ptr(d_workspace), temp_bytes, stream)


def plan(desc: Desc, n: int) -> PlanInfo:
    out = PlanInfo()
    _check(lib().rsort_plan(ctypes.byref(desc), n, ctypes.byref(out)))
    return out


def _ptr(t) -> int | None:
    return None if t is None else int(t.data_ptr())


def _vbytes(values) -> int:
    return 0 if values is None else values.element_size() * (values[0].numel() if values.dim() > 1 else 1)


def _check_tensor(t, name: str, need_bytes: int) -> None:
    """A tensor handed to the library as a device buffer must be a contiguous
    CUDA tensor holding at least need_bytes bytes (the library sees only its
    address, so a smaller, strided or host tensor would be written out of
    bounds)."""
    if t is None:
        return
    if not getattr(t, "is_cuda", False):
        raise ValueError(f"{name} must be a CUDA (device) tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    have = t.numel() * t.element_size()
    if have < need_bytes:
        raise ValueError(f"{name} holds {have} bytes, {need_bytes} needed")


def _sort(keys, values, temp_bytes, workspace, n, stream, descending, key_type, flags):
    kt = key_type or key_type_of(keys.dtype)
    vb = _vbytes(values)
    desc = make_desc(kt.kind, kt.nbytes, vb, descending, flags)
    n = int(n)
    if workspace is not None:
        if temp_bytes is None or int(temp_bytes) < 0:
            raise ValueError("temp_bytes must be the workspace size")
        _check_tensor(keys, "keys", n * kt.nbytes)
        _check_tensor(values, "values", n * vb)
        _check_tensor(workspace, "workspace", int(temp_bytes))
    return sort_raw(desc, _ptr(keys), _ptr(values), _ptr(workspace), temp_bytes, n, stream)


def sort(d_inout, temp_bytes, d_workspace, n, stream=None, key_type: KeyType | None = None, flags: int = 0) -> int:
    """rsort::sort (rsort.cuh:13-30).  d_workspace None -> returns the workspace size."""
    return _sort(d_inout, None, temp_bytes, d_workspace, n, stream, False, key_type, flags)


def sort_descending(d_inout, temp_bytes, d_workspace, n, stream=None, key_type: KeyType | None = None,
                    flags: int = 0) -> int:
    """rsort::sort_descending (rsort.cuh:33-50)."""
    return _sort(d_inout, None, temp_bytes, d_workspace, n, stream, True, key_type, flags)


def sort_pairs(d_inout, d_inout_values, temp_bytes, d_workspace, n, stream=None, key_type: KeyType | None = None,
               flags: int = 0) -> int:
    """rsort::sort_pairs (rsort.cuh:53-71)."""
    return _sort(d_inout, d_inout_values, temp_bytes, d_workspace, n, stream, False, key_type, flags)


def sort_pairs_descending(d_inout, d_inout_values, temp_bytes, d_workspace, n, stream=None,
                          key_type: KeyType | None = None, flags: int = 0) -> int:
    """rsort::sort_pairs_descending (rsort.cuh:74-92)."""
    return _sort(d_inout, d_inout_values, temp_bytes, d_workspace, n, stream, True, key_type, flags)


def sort_profiled(desc: Desc, keys_ptr, vals_ptr, ws_ptr, temp_bytes: int, n: int, stream=None, max_stages: int = 40):
    """rsort_sort_profiled: sorts, synchronises, returns per-stage device ms."""
    stage = (ctypes.c_float * max_stages)()
    cnt = ctypes.c_int(0)
    tb = ctypes.c_size_t(temp_bytes)
    _check(lib().rsort_sort_profiled(ctypes.byref(desc), keys_ptr, vals_ptr, ws_ptr, ctypes.byref(tb), n,
                                     _stream_ptr(stream), stage, max_stages, ctypes.byref(cnt)))
    return [float(stage[i]) for i in range(cnt.value)]


def debug_phase_slots(reset: bool = True) -> list[int]:
    """Per-phase tick sums of a -DRSA_PHASE_TIMING build (empty otherwise)."""
    buf = (ctypes.c_ulonglong * 32)()
    k = lib().rsort_debug_phase_slots(buf, 32, 1 if reset else 0)
    return [int(buf[i]) for i in range(k)]


# ---------------------------------------------------------------------------
# Segmented sort (no reference counterpart: a batch of independent sorts)
# ---------------------------------------------------------------------------
def segmented_sort_raw(desc: Desc, keys_ptr, vals_ptr, n: int, offsets_ptr, offset_bytes: int, num_segments: int,
                       ws_ptr, temp_bytes: int | None, stream=None) -> int:
    """Direct call of rsort_segmented_sort; returns the workspace size (query) or temp_bytes."""
    tb = ctypes.c_size_t(temp_bytes or 0)
    _check(lib().rsort_segmented_sort(ctypes.byref(desc), keys_ptr, vals_ptr, n, offsets_ptr, offset_bytes,
                                      num_segments, ws_ptr, ctypes.byref(tb), _stream_ptr(stream)))
    return int(tb.value)


def segmented_plan(desc: Desc, n: int, offset_bytes: int, num_segments: int) -> SegPlanInfo:
    out = SegPlanInfo()
    _check(lib().rsort_segmented_plan(ctypes.byref(desc), n, offset_bytes, num_segments, ctypes.byref(out)))
    return out


def segmented_sort(keys, offsets, temp_bytes, d_workspace, values=None, descending: bool = False, stream=None,
                   key_type: KeyType | None = None) -> int:
    """Sort every segment keys[offsets[i]:offsets[i+1]] (and values) in place.

    offsets: int32/int64 device tensor of num_segments + 1 entries.
    d_workspace None -> returns the workspace size (no device work).
    """
    kt = key_type or key_type_of(keys.dtype)
    vb = _vbytes(values)
    desc = make_desc(kt.kind, kt.nbytes, vb, descending)
    num_segments = max(int(offsets.numel()) - 1, 0)
    if d_workspace is not None:
        n = int(keys.shape[0])
        if temp_bytes is None or int(temp_bytes) < 0:
            raise ValueError("temp_bytes must be the workspace size")
        _check_tensor(keys, "keys", n * kt.nbytes)
        _check_tensor(values, "values", n * vb)
        _check_tensor(offsets, "offsets", (num_segments + 1) * offsets.element_size())
        _check_tensor(d_workspace, "workspace", int(temp_bytes))
    return segmented_sort_raw(desc, _ptr(keys), _ptr(values), int(keys.shape[0]), _ptr(offsets),
                              offsets.element_size(), num_segments, _ptr(d_workspace), temp_bytes, stream)


# ---------------------------------------------------------------------------
# Serving loop (no reference counterpart): continuous batching of requests
# ---------------------------------------------------------------------------
class Server:
    """Batches sort requests (device tensors, sorted in place) on `depth`
    slots of `batch_keys` keys / `batch_requests` requests.

    mode: "direct" (LDS-bin launches straight on the request buffers),
    "graph" (gather -> segmented sort -> scatter replayed from a hipGraph) or
    "batched" (the same three steps launched directly)."""

    def __init__(self, key_type: KeyType, value_bytes: int = 0, descending: bool = False, batch_keys: int = 1 << 24,
                 batch_requests: int = 4096, window_us: int = 0, depth: int = 2, max_pending: int = 0,
                 mode: str = "direct"):
        cfg = ServeConfig()
        cfg.desc = make_desc(key_type.kind, key_type.nbytes, value_bytes, descending)
        cfg.batch_keys = batch_keys
        cfg.batch_requests = batch_requests
        cfg.window_us = window_us
        cfg.depth = depth
        cfg.max_pending = max_pending
        cfg.flags = {"direct": 0, "graph": SERVE_BATCHED, "batched": SERVE_BATCHED | SERVE_NO_GRAPH}[mode]
        h = ctypes.c_void_p()
        _check(lib().rsort_serve_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.key_bytes = key_type.nbytes
        self.value_bytes = value_bytes

    def submit(self, keys, values=None, n: int | None = None) -> int:
        rid = ctypes.c_uint64(0)
        n = int(keys.shape[0]) if n is None else int(n)
        # tensors are checked for room; any other object with data_ptr() is a
        # raw device address the caller vouches for
        for t, name, b in ((keys, "keys", self.key_bytes), (values, "values", self.value_bytes)):
            if hasattr(t, "numel"):
                _check_tensor(t, name, n * b)
        _check(lib().rsort_serve_submit(self._h, _ptr(keys), _ptr(values), n, ctypes.byref(rid)))
        return int(rid.value)

    def submit_many(self, key_ptrs, val_ptrs, sizes) -> list[int]:
        """Submits requests given as device addresses and sizes in one call."""
        import numpy as np

        count = len(sizes)
        pv = ctypes.POINTER(ctypes.c_void_p)
        pu = ctypes.POINTER(ctypes.c_uint64)
        kp = np.ascontiguousarray(key_ptrs, dtype=np.uint64)
        vp_ = np.ascontiguousarray(val_ptrs, dtype=np.uint64) if val_ptrs is not None else None
        nn = np.ascontiguousarray(sizes, dtype=np.uint64)
        if kp.size != count or nn.size != count or (vp_ is not None and vp_.size != count):
            raise ValueError("key_ptrs, val_ptrs and sizes must have one entry per request")
        ids = np.zeros(count, dtype=np.uint64)
        _check(lib().rsort_serve_submit_many(self._h, kp.ctypes.data_as(pv),
                                             vp_.ctypes.data_as(pv) if vp_ is not None else None,
                                             nn.ctypes.data_as(pu), count, ids.ctypes.data_as(pu)))
        return ids.tolist()

    def wait(self, request_id: int, timeout_us: int = 10_000_000) -> RequestInfo:
        info = RequestInfo()
        _check(lib().rsort_serve_wait(self._h, request_id, timeout_us, ctypes.byref(info)))
        return info

    def wait_many(self, request_ids, timeout_us: int = 10_000_000):
        """Waits for all the requests in one call and releases them; returns
        their RequestInfo records (a ctypes array, in the order given)."""
        import numpy as np

        ids = np.ascontiguousarray(request_ids, dtype=np.uint64)
        infos = (RequestInfo * ids.size)()
        _check(lib().rsort_serve_wait_many(self._h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ids.size,
                                           timeout_us, infos))
        return infos

    def flush(self) -> None:
        _check(lib().rsort_serve_flush(self._h))

    def stats(self) -> ServeStats:
        out = ServeStats()
        _check(lib().rsort_serve_stats(self._h, ctypes.byref(out)))
        return out

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, None
            _check(lib().rsort_serve_destroy(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------
# Tuning table (what autotune.py chooses and persists)
# ---------------------------------------------------------------------------
def tuning_get(key_bytes: int, value_bytes: int) -> Tuning:
    t = Tuning()
    _check(lib().rsort_tuning_get(key_bytes, value_bytes, ctypes.byref(t)))
    return t


def tuning_set(key_bytes: int, value_bytes: int, tile_candidate: int, hist_blocks_per_cu: int,
               lds_threshold: int, small_tile_candidate: int = 0, small_log2n: int = 0,
               tiny_tile_candidate: int = 0, tiny_log2n: int = 0) -> None:
    """Sets one entry; small_log2n = tiny_log2n = 0 uses tile_candidate for every n."""
    t = Tuning(tile_candidate, hist_blocks_per_cu, lds_threshold, small_tile_candidate, small_log2n,
               tiny_tile_candidate, tiny_log2n, 0)
    _check(lib().rsort_tuning_set(key_bytes, value_bytes, ctypes.byref(t)))


def tuning_load(path: str | None) -> int:
    """Reset to the built-in defaults, then read `path` (None: the persisted
    table found at first use; "": built-in defaults only)."""
    return int(lib().rsort_tuning_load(None if path is None else path.encode()))


# ---------------------------------------------------------------------------
# Device generator and checks (reference utils.cuh:227-437 semantics)
# ---------------------------------------------------------------------------
def init_keys(keys, key_bytes: int, mode: int, seed: int, n: int | None = None, stream=None) -> None:
    """Fill a device tensor with the reference generator's keys (mode 1 random, 2 byte_skip, 3 ascending)."""
    n = keys.numel() * keys.element_size() // key_bytes if n is None else n
    _check(lib().rsort_init_keys(key_bytes, mode, seed, _ptr(keys), n, _stream_ptr(stream)))


def set_seed_radix(index: int) -> int:
    return int(lib().rsort_set_seed_radix(index))


def _counter():
    import torch

    return torch.zeros(1, dtype=torch.int64, device="cuda")


def check_sorted(keys, desc: Desc, n: int, stream=None) -> int:
    """Number of adjacent pairs out of order (0 = sorted)."""
    c = _counter()
    _check(lib().rsort_check_sorted(ctypes.byref(desc), _ptr(keys), n, c.data_ptr(), _stream_ptr(stream)))
    return int(c.item())


def check_pairings(keys_orig, keys_sorted, vals_sorted, key_bytes: int, value_bytes: int, n: int, stream=None) -> int:
    """Number of positions violating the value = original index pairing or stability (0 = correct)."""
    c = _counter()
    _check(lib().rsort_check_pairings(key_bytes, value_bytes, _ptr(keys_orig), _ptr(keys_sorted), _ptr(vals_sorted),
                                      n, c.data_ptr(), _stream_ptr(stream)))
    return int(c.item())


def digit_counts(keys, desc: Desc, n: int, stream=None):
    """Per-pass digit counts of the twiddled keys as a [key_bytes, 256] int64 device tensor."""
    import torch

    out = torch.zeros(desc.key_bytes * 256, dtype=torch.int64, device="cuda")
    _check(lib().rsort_digit_counts(ctypes.byref(desc), _ptr(keys), n, out.data_ptr(), _stream_ptr(stream)))
    return out.view(desc.key_bytes, 256)
