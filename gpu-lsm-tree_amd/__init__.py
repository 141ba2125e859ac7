"""Python binding of libgdict (the MI355X dictionaries) over its C ABI.

The classes mirror the reference's host interfaces so that tests read like
the reference's own drivers:

* ``GPULSMOpt`` mirrors ``class GPULSMOpt`` (GPULSMOpt.cuh:5599-5872) with
  the batch structs of gpu_dictionary_adapter.cuh flattened into arguments;
* ``FlixIndex`` mirrors the FliX updatable index
  (FliX/impl_cg_rtx_index_updates.cuh:1711-2693, non-OptiX members).

Arguments are device buffers given as torch tensors (any 4- or 8-byte
integer dtype; the bits are passed through) or as raw integer addresses.
Streams are torch streams, raw ``hipStream_t`` integers, or None for the
current torch stream.  Error codes of the C ABI become the Python
counterparts of the reference's C++ exceptions.

The product path is the HIP library only: if ``libgdict.so`` is missing the
import of this module fails loudly; nothing falls back to a CPU path.
"""

import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgdict.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

GD_OK = 0
GD_INVALID_ARGUMENT = 1
GD_OUT_OF_MEMORY = 2
GD_LENGTH = 3
GD_PUBLICATION_FAILED = 4
GD_HIP_ERROR = 5
GD_CAPACITY = 6

PUBLICATION_OUTPUT_OVERFLOW = 1 << 3
NOT_FOUND = 0xFFFFFFFF


class GdictError(RuntimeError):
    """std::runtime_error (HIP errors, failed publication)."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class InvalidArgument(GdictError, ValueError):
    """std::invalid_argument."""


class OutOfMemory(GdictError, MemoryError):
    """std::bad_alloc."""


class LengthError(GdictError, OverflowError):
    """std::length_error."""


class PublicationFailed(GdictError):
    """std::runtime_error raised by updates after a failed publication."""


class CapacityExhausted(GdictError):
    """Allocation region or memory budget exhausted."""


_ERRORS = {
    GD_INVALID_ARGUMENT: InvalidArgument,
    GD_OUT_OF_MEMORY: OutOfMemory,
    GD_LENGTH: LengthError,
    GD_PUBLICATION_FAILED: PublicationFailed,
    GD_HIP_ERROR: GdictError,
    GD_CAPACITY: CapacityExhausted,
}


class _Config(ctypes.Structure):
    _fields_ = [
        ("max_elements", ctypes.c_uint64),
        ("batch_capacity", ctypes.c_uint64),
        ("level_zero_capacity", ctypes.c_uint64),
        ("delete_order", ctypes.c_uint32),
        ("base_delete_values", ctypes.c_uint32),
    ]


class _FlixConfig(ctypes.Structure):
    _fields_ = [
        ("key_bytes", ctypes.c_uint32),
        ("value_bytes", ctypes.c_uint32),
        ("node_slots", ctypes.c_uint32),
        ("fill_percent", ctypes.c_uint32),
        ("rebuild_chain_nodes", ctypes.c_uint32),
    ]


class _Tuning(ctypes.Structure):
    _fields_ = [
        ("dir_rows_log", ctypes.c_uint32),
        ("merge_tile_log", ctypes.c_uint32),
        ("sort_block", ctypes.c_uint32),
        ("grouped_lookup_min_log", ctypes.c_uint32),
        ("small_batch_max", ctypes.c_uint32),
        ("max_runs", ctypes.c_uint32),
        ("segmented_admission", ctypes.c_uint32),
        ("merge_block", ctypes.c_uint32),
        ("merge_ways", ctypes.c_uint32),
        ("merge_fanout", ctypes.c_uint32),
    ]


TUNING_FIELDS = ("dir_rows_log", "merge_tile_log", "sort_block",
                 "grouped_lookup_min_log", "small_batch_max", "max_runs",
                 "segmented_admission", "merge_block", "merge_ways",
                 "merge_fanout")


def _tuning(d):
    t = _Tuning()
    for k, v in (d or {}).items():
        if k not in TUNING_FIELDS:
            raise InvalidArgument(GD_INVALID_ARGUMENT, "unknown knob " + k)
        setattr(t, k, int(v))
    return t


def tuned_config(section="lsm"):
    """The autotuner's choices (tuned.json next to this file), or {}."""
    import json
    path = os.path.join(_HERE, "tuned.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get(section, {})


class _ServeConfig(ctypes.Structure):
    _fields_ = [
        ("lsm", _Config),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("max_step_keys", ctypes.c_uint64),
        ("max_inflight_steps", ctypes.c_uint32),
        ("key_bytes", ctypes.c_uint32),
        ("tuning", ctypes.POINTER(_Tuning)),
        ("step_launch", ctypes.c_uint32),
    ]


class _ServeStats(ctypes.Structure):
    _fields_ = [
        ("requests", ctypes.c_uint64),
        ("steps", ctypes.c_uint64),
        ("coalesced_updates", ctypes.c_uint64),
        ("coalesced_queries", ctypes.c_uint64),
        ("latency_p50_us", ctypes.c_double),
        ("latency_p99_us", ctypes.c_double),
        ("latency_p999_us", ctypes.c_double),
        ("latency_max_us", ctypes.c_double),
        ("peak_staged_bytes", ctypes.c_uint64),
        ("resident_bytes", ctypes.c_uint64),
        ("graph_steps", ctypes.c_uint64),
        ("live_bytes", ctypes.c_uint64),
        ("peak_resident_bytes", ctypes.c_uint64),
    ]


class _Profile(ctypes.Structure):
    _fields_ = [
        ("merge_launches", ctypes.c_uint64),
        ("merge_ms", ctypes.c_double),
        ("merge_rows_in", ctypes.c_uint64),
        ("merge_rows_out", ctypes.c_uint64),
        ("lookup_launches", ctypes.c_uint64),
        ("lookup_ms", ctypes.c_double),
        ("lookup_queries", ctypes.c_uint64),
        ("admission_batches", ctypes.c_uint64),
        ("admission_ms", ctypes.c_double),
        ("admission_keys", ctypes.c_uint64),
        ("range_launches", ctypes.c_uint64),
        ("range_ms", ctypes.c_double),
        ("range_queries", ctypes.c_uint64),
        ("range_runs", ctypes.c_uint64),
    ]


class _Memory(ctypes.Structure):
    _fields_ = [
        ("held_bytes", ctypes.c_uint64),
        ("live_bytes", ctypes.c_uint64),
        ("cached_bytes", ctypes.c_uint64),
        ("peak_held_bytes", ctypes.c_uint64),
        ("trimmed_blocks", ctypes.c_uint64),
        ("row_bytes_ub", ctypes.c_uint64),
    ]


def _memory_dict(m):
    return {name: int(getattr(m, name)) for name, _ in m._fields_}


_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_I = ctypes.c_int

# symbol -> argtypes (restype is gd_status = int unless noted)
SIGNATURES = {
    "gd_lsm_create": [ctypes.POINTER(_Config), ctypes.POINTER(_P)],
    "gd_lsm_destroy": [_P],
    "gd_lsm_clear": [_P, _P],
    "gd_lsm_bulk_build": [_P, _P, _P, _U64, _P],
    "gd_lsm_insert": [_P, _P, _P, _U64, _P],
    "gd_lsm_erase": [_P, _P, _U64, _P],
    "gd_lsm_apply_mixed": [_P, _P, _P, _P, _U64, _P],
    "gd_lsm_lookup": [_P, _P, _U64, _P, _P, _I, _P],
    "gd_lsm_lookup_noalloc": [_P, _P, _U64, _P, _P, _P],
    "gd_lsm_range_sum": [_P, _P, _P, _U64, _P, _P],
    "gd_lsm_range_count": [_P, _P, _P, _U64, _P, _P],
    "gd_lsm_successor": [_P, _P, _U64, _P, _P],
    "gd_lsm_dump_sorted": [_P, _P, _P, _U64, ctypes.POINTER(_U64), _P],
    "gd_lsm_compact": [_P, _P],
    "gd_lsm_carry_status": [_P, ctypes.POINTER(ctypes.c_uint32)],
    "gd_lsm_resident_bytes": [_P, ctypes.POINTER(_U64)],
    "gd_lsm_memory": [_P, ctypes.POINTER(_Memory)],
    "gd_lsm_trim": [_P],
    "gd_lsm_shape": [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(_U64)],
    "gd_lsm_set_tuning": [_P, ctypes.POINTER(_Tuning)],
    "gd_lsm_set_profiling": [_P, _I],
    "gd_lsm_read_profile": [_P, ctypes.POINTER(_Profile), _I],
    "gd_lsm_get_tuning": [_P, ctypes.POINTER(_Tuning)],
    "gd_lsm_snapshot_save": [_P, ctypes.c_char_p, ctypes.POINTER(_U64), _P],
    "gd_lsm_snapshot_restore": [_P, ctypes.c_char_p, ctypes.POINTER(_U64),
                                _P],
    "gd_lsm64_snapshot_save": [_P, ctypes.c_char_p, ctypes.POINTER(_U64),
                               _P],
    "gd_lsm64_snapshot_restore": [_P, ctypes.c_char_p, ctypes.POINTER(_U64),
                                  _P],
    "gd_lsm64_create": [ctypes.POINTER(_Config), ctypes.POINTER(_P)],
    "gd_lsm64_destroy": [_P],
    "gd_lsm64_clear": [_P, _P],
    "gd_lsm64_bulk_build": [_P, _P, _P, _U64, _P],
    "gd_lsm64_insert": [_P, _P, _P, _U64, _P],
    "gd_lsm64_erase": [_P, _P, _U64, _P],
    "gd_lsm64_apply_mixed": [_P, _P, _P, _P, _U64, _P],
    "gd_lsm64_lookup": [_P, _P, _U64, _P, _P, _P],
    "gd_lsm64_range_sum": [_P, _P, _P, _U64, _P, _P],
    "gd_lsm64_range_count": [_P, _P, _P, _U64, _P, _P],
    "gd_lsm64_successor": [_P, _P, _U64, _P, _P],
    "gd_lsm64_dump_sorted": [_P, _P, _P, _U64, ctypes.POINTER(_U64), _P],
    "gd_lsm64_compact": [_P, _P],
    "gd_lsm64_carry_status": [_P, ctypes.POINTER(ctypes.c_uint32)],
    "gd_lsm64_resident_bytes": [_P, ctypes.POINTER(_U64)],
    "gd_lsm64_memory": [_P, ctypes.POINTER(_Memory)],
    "gd_lsm64_trim": [_P],
    "gd_lsm64_shape": [_P, ctypes.POINTER(ctypes.c_uint32),
                       ctypes.POINTER(_U64)],
    "gd_lsm64_set_tuning": [_P, ctypes.POINTER(_Tuning)],
    "gd_lsm64_get_tuning": [_P, ctypes.POINTER(_Tuning)],
    "gd_serve_create": [ctypes.POINTER(_ServeConfig), ctypes.POINTER(_P)],
    "gd_serve_destroy": [_P],
    "gd_serve_submit": [_P, ctypes.c_int, _P, _P, _U64, _P, _P,
                        ctypes.POINTER(_U64)],
    "gd_serve_wait": [_P, _U64, ctypes.POINTER(_U64)],
    "gd_serve_drain": [_P],
    "gd_serve_get_stats": [_P, ctypes.POINTER(_ServeStats), _I],
    "gd_flix_create": [ctypes.POINTER(_FlixConfig), ctypes.POINTER(_P)],
    "gd_flix_destroy": [_P],
    "gd_flix_build": [_P, _P, _U64, _U64, _P],
    "gd_flix_insert": [_P, _P, _P, _U64, _P],
    "gd_flix_remove": [_P, _P, _U64, _P],
    "gd_flix_remove_sorted": [_P, _P, _U64, _P],
    "gd_flix_lookup_sorted": [_P, _P, _U64, _P, _P],
    "gd_flix_lookup": [_P, _P, _U64, _P, _P],
    "gd_flix_successor_sorted": [_P, _P, _U64, _P, _P],
    "gd_flix_rebuild": [_P, _P],
    "gd_flix_range": [_P, _P, _P, _U64, _P, _P, _P, _P],
    "gd_flix_total_keys": [_P, ctypes.POINTER(_U64), _P],
    "gd_flix_resident_bytes": [_P, ctypes.POINTER(_U64)],
    "gd_flix_memory": [_P, ctypes.POINTER(_Memory)],
    "gd_flix_trim": [_P],
}

_lib = None


def build(verbose=False):
    """Compiles libgdict.so in-tree for gfx950 (hipcc, no GPU needed).
    Always recompiles (make -B), so a call is a real build check."""
    cmd = ["make", "-B", "-s", "-j8", "-C", _HERE]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("building libgdict.so failed:\n" + res.stdout +
                           res.stderr)
    if verbose:
        print(res.stdout)


def lib():
    """Loads libgdict.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libgdict.so is missing at %s; run build() first" % LIB_PATH)
    handle = ctypes.CDLL(LIB_PATH)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(handle, name, None)
        if fn is None:
            continue
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    handle.gd_last_error_message.argtypes = [_P]
    handle.gd_last_error_message.restype = ctypes.c_char_p
    _lib = handle
    return _lib


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _stream(s):
    if s is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(s, int):
        return s
    return s.cuda_stream


def _bad(msg):
    return InvalidArgument(GD_INVALID_ARGUMENT, msg)


def _count(x, count):
    """The element count of a call: `count` when given, else x.numel().
    An explicit count beyond a tensor's length is rejected."""
    if count is not None:
        n = int(count)
        if n < 0:
            raise _bad("negative count")
        if x is not None and not isinstance(x, int) and n > x.numel():
            raise _bad("count %d exceeds the %d elements of the buffer"
                       % (n, x.numel()))
        return n
    if x is None:
        return 0
    if isinstance(x, int):
        raise _bad("a raw device address needs an explicit count")
    return int(x.numel())


def check_buffer(x, name, width, need, optional=False):
    """Validates one device-buffer argument before anything is launched and
    returns its address.

    A tensor must live on a GPU, be contiguous, have `width`-byte elements
    and hold at least `need` of them; None passes only when `optional` (or
    when nothing is read or written, need == 0).  Raw integer addresses are
    the caller's responsibility and pass unchecked.  Violations raise
    InvalidArgument, so no kernel ever reads or writes past a tensor."""
    if x is None:
        if optional or need == 0:
            return None
        raise _bad("%s is required" % name)
    if isinstance(x, int):
        return x
    if not getattr(x, "is_cuda", False):
        raise _bad("%s must be a GPU tensor" % name)
    if not x.is_contiguous():
        raise _bad("%s must be contiguous" % name)
    if x.element_size() != width:
        raise _bad("%s has %d-byte elements, %d-byte ones are required"
                   % (name, x.element_size(), width))
    if x.numel() < need:
        raise _bad("%s holds %d elements, %d are required"
                   % (name, x.numel(), need))
    return x.data_ptr()


def _check(status, handle):
    if status == GD_OK:
        return
    msg = lib().gd_last_error_message(handle)
    msg = msg.decode() if msg else ""
    raise _ERRORS.get(status, GdictError)(status, msg)


class GPULSMOpt:
    """Mirror of GPULSMOpt (GPULSMOpt.cuh:5599) over gd_lsm_*; key_bytes=8
    selects the uint64 key/value dictionary (gd_lsm64_*)."""

    def __init__(self, max_elements=0, batch_capacity=0,
                 level_zero_capacity=0, tuning=None, key_bytes=4):
        """tuning: dict of gd_lsm_tuning knobs, or "auto" for tuned.json."""
        if key_bytes not in (4, 8):
            raise InvalidArgument(GD_INVALID_ARGUMENT,
                                  "key_bytes must be 4 or 8")
        self.key_bytes = key_bytes
        self._p = "gd_lsm_" if key_bytes == 4 else "gd_lsm64_"
        cfg = _Config(max_elements, batch_capacity, level_zero_capacity, 0, 0)
        h = _P()
        _check(getattr(lib(), self._p + "create")(ctypes.byref(cfg),
                                                   ctypes.byref(h)), None)
        self._h = h
        if tuning == "auto":
            tuning = tuned_config("lsm")
        if tuning:
            self.set_tuning(tuning)

    def set_tuning(self, knobs):
        t = _tuning(knobs)
        self._call("set_tuning", ctypes.byref(t))

    def set_profiling(self, enable=True):
        self._call("set_profiling", 1 if enable else 0)

    def read_profile(self, reset=True):
        p = _Profile()
        self._call("read_profile", ctypes.byref(p), 1 if reset else 0)
        return {f: getattr(p, f) for f, _ in _Profile._fields_}

    def get_tuning(self):
        t = _Tuning()
        self._call("get_tuning", ctypes.byref(t))
        return {k: int(getattr(t, k)) for k in TUNING_FIELDS}

    def close(self):
        if getattr(self, "_h", None):
            getattr(lib(), self._p + "destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        _check(getattr(lib(), self._p + name)(self._h, *args), self._h)

    def clear(self, stream=None):
        self._call("clear", _stream(stream))

    def _keys(self, x, name, n):
        return check_buffer(x, name, self.key_bytes, n)

    def bulk_build(self, keys, values, count=None, stream=None):
        n = _count(keys, count)
        self._call("bulk_build", self._keys(keys, "keys", n),
                   self._keys(values, "values", n), n, _stream(stream))

    def insert(self, keys, values, count=None, stream=None):
        n = _count(keys, count)
        self._call("insert", self._keys(keys, "keys", n),
                   self._keys(values, "values", n), n, _stream(stream))

    def erase(self, keys, count=None, stream=None):
        n = _count(keys, count)
        self._call("erase", self._keys(keys, "keys", n), n, _stream(stream))

    def apply_mixed(self, keys, values, is_erase, count=None, stream=None):
        n = _count(keys, count)
        self._call("apply_mixed", self._keys(keys, "keys", n),
                   self._keys(values, "values", n),
                   check_buffer(is_erase, "is_erase", 1, n), n,
                   _stream(stream))

    def lookup(self, queries, out_values, out_found=None, count=None,
               stream=None, quotients_grouped=False):
        n = _count(queries, count)
        q = self._keys(queries, "queries", n)
        ov = self._keys(out_values, "out_values", n)
        of = check_buffer(out_found, "out_found", 1, n, optional=True)
        if self.key_bytes == 4:
            self._call("lookup", q, n, ov, of, 1 if quotients_grouped else 0,
                       _stream(stream))
        else:
            self._call("lookup", q, n, ov, of, _stream(stream))

    def lookup_noalloc(self, queries, out_values, out_found=None, count=None,
                       stream=None):
        """gd_lsm_lookup_noalloc: lookup() without any device allocation
        (uint32 dictionaries, at most 2^20 queries)."""
        if self.key_bytes != 4:
            raise _bad("lookup_noalloc serves uint32 dictionaries")
        n = _count(queries, count)
        q = self._keys(queries, "queries", n)
        ov = self._keys(out_values, "out_values", n)
        of = check_buffer(out_found, "out_found", 1, n, optional=True)
        self._call("lookup_noalloc", q, n, ov, of, _stream(stream))

    def range(self, lo, hi, out_sums, count=None, stream=None):
        n = _count(lo, count)
        self._call("range_sum", self._keys(lo, "lo", n),
                   self._keys(hi, "hi", n), n,
                   self._keys(out_sums, "out_sums", n), _stream(stream))

    def range_count(self, lo, hi, out_counts, count=None, stream=None):
        n = _count(lo, count)
        self._call("range_count", self._keys(lo, "lo", n),
                   self._keys(hi, "hi", n), n,
                   self._keys(out_counts, "out_counts", n), _stream(stream))

    def successor(self, queries, out_keys, count=None, stream=None):
        n = _count(queries, count)
        self._call("successor", self._keys(queries, "queries", n), n,
                   self._keys(out_keys, "out_keys", n), _stream(stream))

    def dump_sorted(self, out_keys=None, out_values=None, capacity=0,
                    stream=None):
        """Returns the live count; fills out_keys/out_values when given."""
        capacity = int(capacity)
        ok = check_buffer(out_keys, "out_keys", self.key_bytes, capacity,
                          optional=True)
        ov = check_buffer(out_values, "out_values", self.key_bytes, capacity,
                          optional=True)
        n = _U64(0)
        st = getattr(lib(), self._p + "dump_sorted")(
            self._h, ok, ov, capacity, ctypes.byref(n), _stream(stream))
        if st not in (GD_OK, GD_LENGTH):
            _check(st, self._h)
        return int(n.value)

    def compact(self, stream=None):
        self._call("compact", _stream(stream))

    def save(self, path, stream=None):
        """Writes the live contents to a snapshot file; returns the count."""
        n = _U64(0)
        self._call("snapshot_save", os.fsencode(path), ctypes.byref(n),
                   _stream(stream))
        return int(n.value)

    def restore(self, path, stream=None):
        """Replaces the contents with a snapshot's; returns the count."""
        n = _U64(0)
        self._call("snapshot_restore", os.fsencode(path), ctypes.byref(n),
                   _stream(stream))
        return int(n.value)

    def canonical_carry_status(self):
        v = ctypes.c_uint32(0)
        self._call("carry_status", ctypes.byref(v))
        return int(v.value)

    def gpu_resident_bytes(self):
        v = _U64(0)
        self._call("resident_bytes", ctypes.byref(v))
        return int(v.value)

    def memory(self):
        """Held, live, cached and peak device bytes of this handle, the
        blocks trimmed so far and the rows' bytes (upper bound)."""
        m = _Memory()
        self._call("memory", ctypes.byref(m))
        return _memory_dict(m)

    def trim(self):
        """Returns every cached block to HIP (waits for outstanding work)."""
        self._call("trim")

    def shape(self):
        runs = ctypes.c_uint32(0)
        rows = _U64(0)
        self._call("shape", ctypes.byref(runs), ctypes.byref(rows))
        return int(runs.value), int(rows.value)


class Server:
    """Continuous-batching engine over one LSM dictionary (gd_serve_*).

    submit() returns a ticket; the request's device buffers must stay alive
    until wait(ticket) returns (this wrapper keeps references until then).
    """

    INSERT, ERASE, LOOKUP, RANGE_COUNT, RANGE_SUM = range(5)

    def __init__(self, max_elements=0, batch_capacity=0,
                 level_zero_capacity=0, hbm_budget_bytes=0, max_step_keys=0,
                 max_inflight_steps=0, tuning=None, key_bytes=4,
                 step_launch="kernels"):
        """step_launch: "kernels" (every step kernel by kernel), "direct"
        (small steps are fused kernels launched directly) or "graph" (each
        small step is one captured graph launch)."""
        modes = {"kernels": 0, "direct": 1, "graph": 2}
        if step_launch not in modes:
            raise _bad("step_launch must be one of %s" % sorted(modes))
        if tuning == "auto":
            tuning = tuned_config("lsm")
        self._tuning = _tuning(tuning) if tuning else None
        self.key_bytes = key_bytes
        cfg = _ServeConfig(
            _Config(max_elements, batch_capacity, level_zero_capacity, 0, 0),
            hbm_budget_bytes, max_step_keys, max_inflight_steps, key_bytes,
            ctypes.pointer(self._tuning) if self._tuning else None,
            modes[step_launch])
        h = _P()
        _check(lib().gd_serve_create(ctypes.byref(cfg), ctypes.byref(h)),
               None)
        self._h = h
        self._keep = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().gd_serve_destroy(self._h)
            self._h = None
            self._keep = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, op, a, b=None, out=None, out_found=None, count=None):
        n = _count(a, count)
        w = self.key_bytes
        op = int(op)
        pa = check_buffer(a, "a", w, n)
        if op == self.INSERT:
            pb = check_buffer(b, "b (values)", w, n)
        elif op in (self.RANGE_COUNT, self.RANGE_SUM):
            pb = check_buffer(b, "b (hi)", w, n)
        else:
            pb = check_buffer(b, "b", w, n, optional=True)
        is_query = op in (self.LOOKUP, self.RANGE_COUNT, self.RANGE_SUM)
        po = check_buffer(out, "out", w, n if is_query else 0,
                          optional=not is_query)
        pf = check_buffer(out_found, "out_found", 1, n, optional=True)
        t = _U64(0)
        _check(lib().gd_serve_submit(self._h, op, pa, pb, n, po, pf,
                                     ctypes.byref(t)), self._h)
        self._keep[t.value] = (a, b, out, out_found)
        return int(t.value)

    def wait(self, ticket):
        """Returns the step that served the request."""
        step = _U64(0)
        st = lib().gd_serve_wait(self._h, int(ticket), ctypes.byref(step))
        self._keep.pop(int(ticket), None)
        _check(st, self._h)
        return int(step.value)

    def drain(self):
        _check(lib().gd_serve_drain(self._h), self._h)

    def stats(self, reset=False):
        s = _ServeStats()
        _check(lib().gd_serve_get_stats(self._h, ctypes.byref(s),
                                        1 if reset else 0), self._h)
        return {f: getattr(s, f) for f, _ in _ServeStats._fields_}


class FlixIndex:
    """Mirror of the FliX updatable index over gd_flix_*."""

    def __init__(self, key_bytes=4, value_bytes=4, node_slots=0,
                 fill_percent=0, tuning=None, rebuild_chain_nodes=None):
        """tuning="auto": node_slots, fill_percent and (unless given)
        rebuild_chain_nodes from tuned.json.  rebuild_chain_nodes > 0: an
        insert leaving a chain longer than that rebuilds the bucket layer
        afterwards; 0 never rebuilds."""
        if tuning == "auto":
            t = tuned_config("flix")
            node_slots = node_slots or t.get("node_slots", 0)
            fill_percent = fill_percent or t.get("fill_percent", 0)
            if rebuild_chain_nodes is None:
                rebuild_chain_nodes = t.get("rebuild_chain_nodes", 0)
        if rebuild_chain_nodes is None:
            rebuild_chain_nodes = 0
        cfg = _FlixConfig(key_bytes, value_bytes, node_slots, fill_percent,
                          rebuild_chain_nodes)
        h = _P()
        _check(lib().gd_flix_create(ctypes.byref(cfg), ctypes.byref(h)), None)
        self._h = h
        self.key_bytes = key_bytes
        self.value_bytes = value_bytes

    def close(self):
        if getattr(self, "_h", None):
            lib().gd_flix_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        _check(getattr(lib(), name)(self._h, *args), self._h)

    def _k(self, x, name, n, optional=False):
        return check_buffer(x, name, self.key_bytes, n, optional)

    def _v(self, x, name, n):
        return check_buffer(x, name, self.value_bytes, n)

    def build_bucket_layer_only(self, keys, max_size=0, count=None,
                                stream=None):
        n = _count(keys, count)
        self._call("gd_flix_build", self._k(keys, "keys", n), n,
                   int(max_size), _stream(stream))

    def insert(self, keys, values, count=None, stream=None):
        n = _count(keys, count)
        self._call("gd_flix_insert", self._k(keys, "keys", n),
                   self._v(values, "values", n), n, _stream(stream))

    def remove(self, keys, count=None, stream=None, assume_sorted=False):
        """assume_sorted=True: the keys are ascending already (the
        reference's remove(sorted_keys, ...) contract) and are not sorted
        again; unsorted keys then give undefined results."""
        n = _count(keys, count)
        self._call("gd_flix_remove_sorted" if assume_sorted
                   else "gd_flix_remove", self._k(keys, "keys", n), n,
                   _stream(stream))

    def lookups_ordered(self, sorted_keys, out_values, count=None,
                        stream=None):
        n = _count(sorted_keys, count)
        self._call("gd_flix_lookup_sorted", self._k(sorted_keys, "keys", n),
                   n, self._v(out_values, "out_values", n), _stream(stream))

    def lookup(self, keys, out_values, count=None, stream=None):
        n = _count(keys, count)
        self._call("gd_flix_lookup", self._k(keys, "keys", n), n,
                   self._v(out_values, "out_values", n), _stream(stream))

    def lookups_successor(self, sorted_keys, out_keys, count=None,
                          stream=None):
        n = _count(sorted_keys, count)
        self._call("gd_flix_successor_sorted",
                   self._k(sorted_keys, "keys", n), n,
                   self._k(out_keys, "out_keys", n), _stream(stream))

    def rebuild_bucket_layer_only(self, stream=None):
        self._call("gd_flix_rebuild", _stream(stream))

    def range(self, lo, hi, out_offsets, out_keys=None, out_values=None,
              count=None, stream=None):
        """Pass 1 (out_keys None): per-query counts, scanned into
        out_offsets[0..n] (n + 1 uint64).  Pass 2: the pairs of query i at
        out_keys/out_values[out_offsets[i]:out_offsets[i + 1]]; both
        buffers must hold out_offsets[n] entries."""
        n = _count(lo, count)
        po = check_buffer(out_offsets, "out_offsets", 8, n + 1)
        pk = self._k(out_keys, "out_keys", 0, optional=True)
        pv = check_buffer(out_values, "out_values", self.value_bytes, 0,
                          optional=True)
        if (pk is None) != (pv is None):
            raise _bad("out_keys and out_values go together")
        self._call("gd_flix_range", self._k(lo, "lo", n), self._k(hi, "hi", n),
                   n, po, pk, pv, _stream(stream))

    def compute_total_size(self, stream=None):
        v = _U64(0)
        self._call("gd_flix_total_keys", ctypes.byref(v), _stream(stream))
        return int(v.value)

    def gpu_resident_bytes(self):
        v = _U64(0)
        self._call("gd_flix_resident_bytes", ctypes.byref(v))
        return int(v.value)

    def memory(self):
        """As GPULSMOpt.memory(); row_bytes_ub is the node layer's bytes."""
        m = _Memory()
        self._call("gd_flix_memory", ctypes.byref(m))
        return _memory_dict(m)

    def trim(self):
        self._call("gd_flix_trim")
