"""ctypes wrapper of the oracle (oracle.cpp): TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU-baseline leg of bench.py may use
this module, and only as the checker or the CPU baseline.  The product
(libgdict.so and its binding) never imports it.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_lib = None

_SIGS = {
    "orc_splitmix64": ([_U64], _U64),
    "orc_murmur3_u32": ([_U32, _U32], _U32),
    "orc_murmur3_bytes": ([_P, _U64, _U32], _U32),
    "orc_log_uniform": ([_U64, _U64, _U64], _U64),
    "orc_fill_splitmix64": ([_U64, _U64, _P], None),
    "orc_fill_murmur3": ([_P, _U64, _U32, _P], None),
    "orc_fill_log_uniform": ([_U64, _U64, _U64, _U64, _P], None),
    "orc_zipf_fill": ([ctypes.c_double, _U64, _U64, _U64, _P], None),
    "orc_paper_key_for_index": ([_U32], _U32),
    "orc_paper_mix32": ([_U32], _U32),
    "orc_range_digest": ([_P, _U64, _U64, _P, _P], None),
    "orc_lsm_create": ([_U64, _U64, _U64], _P),
    "orc_lsm_destroy": ([_P], None),
    "orc_lsm_clear": ([_P], None),
    "orc_lsm_bulk_build": ([_P, _P, _P, _U64], ctypes.c_int),
    "orc_lsm_insert": ([_P, _P, _P, _U64], ctypes.c_int),
    "orc_lsm_erase": ([_P, _P, _U64], ctypes.c_int),
    "orc_lsm_apply_mixed": ([_P, _P, _P, _P, _U64], ctypes.c_int),
    "orc_lsm_lookup": ([_P, _P, _U64, _P, _P], None),
    "orc_lsm_range_sum": ([_P, _P, _P, _U64, _P], None),
    "orc_lsm_range_count": ([_P, _P, _P, _U64, _P], None),
    "orc_lsm_successor": ([_P, _P, _U64, _P], None),
    "orc_lsm_size": ([_P], _U64),
    "orc_lsm_dump": ([_P, _P, _P], None),
    "orc_lsm_status": ([_P], _U32),
    "orc_lsm_capacity": ([_P], _U64),
    "orc_lsm_message": ([_P], ctypes.c_char_p),
    "orc_lsm64_create": ([_U64, _U64], _P),
    "orc_lsm64_destroy": ([_P], None),
    "orc_lsm64_clear": ([_P], None),
    "orc_lsm64_bulk_build": ([_P, _P, _P, _U64], ctypes.c_int),
    "orc_lsm64_insert": ([_P, _P, _P, _U64], ctypes.c_int),
    "orc_lsm64_erase": ([_P, _P, _U64], ctypes.c_int),
    "orc_lsm64_apply_mixed": ([_P, _P, _P, _P, _U64], ctypes.c_int),
    "orc_lsm64_lookup": ([_P, _P, _U64, _P, _P], None),
    "orc_lsm64_range_sum": ([_P, _P, _P, _U64, _P], None),
    "orc_lsm64_range_count": ([_P, _P, _P, _U64, _P], None),
    "orc_lsm64_successor": ([_P, _P, _U64, _P], None),
    "orc_lsm64_size": ([_P], _U64),
    "orc_lsm64_dump": ([_P, _P, _P], None),
    "orc_lsm64_status": ([_P], _U32),
    "orc_flix_create": ([ctypes.c_int, ctypes.c_int], _P),
    "orc_flix_destroy": ([_P], None),
    "orc_flix_build": ([_P, _P, _U64], None),
    "orc_flix_insert": ([_P, _P, _P, _U64], None),
    "orc_flix_remove": ([_P, _P, _U64], None),
    "orc_flix_lookup": ([_P, _P, _U64, _P], None),
    "orc_flix_successor": ([_P, _P, _U64, _P], None),
    "orc_flix_range_counts": ([_P, _P, _P, _U64, _P], None),
    "orc_flix_range_enumerate": ([_P, _P, _P, _U64, _P, _P, _P], None),
    "orc_flix_size": ([_P], _U64),
    "orc_flix_dump": ([_P, _P, _P], None),
    "cpu_lsm_create": ([ctypes.c_uint], _P),
    "cpu_lsm_destroy": ([_P], None),
    "cpu_lsm_clear": ([_P], None),
    "cpu_lsm_apply": ([_P, _P, _P, _P, _U64, ctypes.c_int], None),
    "cpu_lsm_lookup": ([_P, _P, _U64, _P, _P], None),
    "cpu_lsm_range": ([_P, _P, _P, _U64, _P, ctypes.c_int], None),
    "cpu_lsm_rows": ([_P], _U64),
    "cpu_lsm_runs": ([_P], ctypes.c_uint),
}


def build():
    res = subprocess.run(["make", "-s", "-C", _HERE], capture_output=True,
                         text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = res
        _lib = h
    return _lib


def _a(x):
    return None if x is None else x.ctypes.data


def u32(x):
    return np.ascontiguousarray(x, dtype=np.uint32)


# --------------------------------------------------------------- generators
def splitmix64_stream(seed, n):
    out = np.empty(n, dtype=np.uint64)
    lib().orc_fill_splitmix64(seed, n, _a(out))
    return out


def murmur3(keys, seed=0):
    keys = u32(keys)
    out = np.empty(len(keys), dtype=np.uint32)
    lib().orc_fill_murmur3(_a(keys), len(keys), seed, _a(out))
    return out


def log_uniform_stream(seed, n, a, b):
    out = np.empty(n, dtype=np.uint64)
    lib().orc_fill_log_uniform(seed, n, a, b, _a(out))
    return out


def zipf_ranks(s, universe, seed, n):
    out = np.empty(n, dtype=np.uint64)
    lib().orc_zipf_fill(s, universe, seed, n, _a(out))
    return out


def range_digest(results, query_offset=0):
    r = u32(results)
    s = _U64(0)
    x = _U64(0)
    lib().orc_range_digest(_a(r), len(r), query_offset, ctypes.byref(s),
                           ctypes.byref(x))
    return int(s.value), int(x.value)


# ---------------------------------------------------------------- LSM oracle
class LsmOracle:
    """Replays GPULSMOpt's observable semantics (see oracle.cpp).  key_bytes
    = 8 selects the 64-bit extension (uint64 keys and values)."""

    def __init__(self, max_elements=0, batch_capacity=0,
                 level_zero_capacity=0, key_bytes=4):
        self.key_bytes = key_bytes
        self.kt = np.uint32 if key_bytes == 4 else np.uint64
        self._p = "orc_lsm_" if key_bytes == 4 else "orc_lsm64_"
        if key_bytes == 4:
            self._h = lib().orc_lsm_create(max_elements, batch_capacity,
                                           level_zero_capacity)
        else:
            self._h = lib().orc_lsm64_create(max_elements, batch_capacity)
        if not self._h:
            raise ValueError("GPULSMOpt capacity exceeds 32-bit key space")

    def _f(self, name):
        return getattr(lib(), self._p + name)

    def _k(self, a):
        return np.ascontiguousarray(a, dtype=self.kt)

    def __del__(self):
        if getattr(self, "_h", None):
            self._f("destroy")(self._h)
            self._h = None

    def clear(self):
        self._f("clear")(self._h)

    def bulk_build(self, keys, values):
        k, v = self._k(keys), self._k(values)
        return self._f("bulk_build")(self._h, _a(k), _a(v), len(k))

    def insert(self, keys, values):
        k, v = self._k(keys), self._k(values)
        return self._f("insert")(self._h, _a(k), _a(v), len(k))

    def erase(self, keys):
        k = self._k(keys)
        return self._f("erase")(self._h, _a(k), len(k))

    def apply_mixed(self, keys, values, is_erase):
        k, v = self._k(keys), self._k(values)
        e = np.ascontiguousarray(is_erase, dtype=np.uint8)
        return self._f("apply_mixed")(self._h, _a(k), _a(v), _a(e), len(k))

    def lookup(self, queries, with_found=False):
        q = self._k(queries)
        out = np.empty(len(q), dtype=self.kt)
        found = np.empty(len(q), dtype=np.uint8) if with_found else None
        self._f("lookup")(self._h, _a(q), len(q), _a(out), _a(found))
        return (out, found) if with_found else out

    def range_sum(self, lo, hi):
        lo, hi = self._k(lo), self._k(hi)
        out = np.empty(len(lo), dtype=self.kt)
        self._f("range_sum")(self._h, _a(lo), _a(hi), len(lo), _a(out))
        return out

    def range_count(self, lo, hi):
        lo, hi = self._k(lo), self._k(hi)
        out = np.empty(len(lo), dtype=self.kt)
        self._f("range_count")(self._h, _a(lo), _a(hi), len(lo), _a(out))
        return out

    def successor(self, queries):
        q = self._k(queries)
        out = np.empty(len(q), dtype=self.kt)
        self._f("successor")(self._h, _a(q), len(q), _a(out))
        return out

    def size(self):
        return int(self._f("size")(self._h))

    def dump(self):
        n = self.size()
        k = np.empty(n, dtype=self.kt)
        v = np.empty(n, dtype=self.kt)
        if n:
            self._f("dump")(self._h, _a(k), _a(v))
        return k, v

    def status(self):
        return int(self._f("status")(self._h))

    def capacity(self):
        return int(lib().orc_lsm_capacity(self._h)) if self.key_bytes == 4 \
            else None

    def message(self):
        return lib().orc_lsm_message(self._h).decode() \
            if self.key_bytes == 4 else ""


# --------------------------------------------------------------- FliX oracle
class FlixOracle:
    def __init__(self, key_bytes=4, value_bytes=4):
        self.kt = np.uint32 if key_bytes == 4 else np.uint64
        self.vt = np.uint32 if value_bytes == 4 else np.uint64
        self._h = lib().orc_flix_create(key_bytes, value_bytes)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_flix_destroy(self._h)
            self._h = None

    def _k(self, x):
        return np.ascontiguousarray(x, dtype=self.kt)

    def build(self, keys):
        k = self._k(keys)
        lib().orc_flix_build(self._h, _a(k), len(k))

    def insert(self, keys, values):
        k = self._k(keys)
        v = np.ascontiguousarray(values, dtype=self.vt)
        lib().orc_flix_insert(self._h, _a(k), _a(v), len(k))

    def remove(self, keys):
        k = self._k(keys)
        lib().orc_flix_remove(self._h, _a(k), len(k))

    def lookup(self, queries):
        q = self._k(queries)
        out = np.empty(len(q), dtype=self.vt)
        lib().orc_flix_lookup(self._h, _a(q), len(q), _a(out))
        return out

    def successor(self, queries):
        q = self._k(queries)
        out = np.empty(len(q), dtype=self.kt)
        lib().orc_flix_successor(self._h, _a(q), len(q), _a(out))
        return out

    def range(self, lo, hi):
        lo, hi = self._k(lo), self._k(hi)
        counts = np.empty(len(lo), dtype=np.uint64)
        lib().orc_flix_range_counts(self._h, _a(lo), _a(hi), len(lo),
                                    _a(counts))
        offsets = np.zeros(len(lo) + 1, dtype=np.uint64)
        np.cumsum(counts, out=offsets[1:])
        total = int(offsets[-1])
        ok = np.empty(total, dtype=self.kt)
        ov = np.empty(total, dtype=self.vt)
        lib().orc_flix_range_enumerate(self._h, _a(lo), _a(hi), len(lo),
                                       _a(offsets), _a(ok), _a(ov))
        return offsets, ok, ov

    def size(self):
        return int(lib().orc_flix_size(self._h))

    def dump(self):
        n = self.size()
        k = np.empty(n, dtype=self.kt)
        v = np.empty(n, dtype=self.vt)
        if n:
            lib().orc_flix_dump(self._h, _a(k), _a(v))
        return k, v


# ------------------------------------------------------- CPU LSM baseline
class CpuLsm:
    """Multithreaded sorted-array LSM (cpu_lsm.cpp), the CPU baseline."""

    def __init__(self, threads):
        self.threads = threads
        self._h = lib().cpu_lsm_create(threads)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().cpu_lsm_destroy(self._h)
            self._h = None

    def clear(self):
        lib().cpu_lsm_clear(self._h)

    def apply(self, keys, values=None, is_erase=None, mode=None):
        k = u32(keys)
        v = u32(values) if values is not None else None
        e = (np.ascontiguousarray(is_erase, dtype=np.uint8)
             if is_erase is not None else None)
        if mode is None:
            mode = 2 if e is not None else (0 if v is not None else 1)
        lib().cpu_lsm_apply(self._h, _a(k), _a(v), _a(e), len(k), mode)

    def lookup(self, queries, with_found=False):
        q = u32(queries)
        out = np.empty(len(q), dtype=np.uint32)
        found = np.empty(len(q), dtype=np.uint8) if with_found else None
        lib().cpu_lsm_lookup(self._h, _a(q), len(q), _a(out), _a(found))
        return (out, found) if with_found else out

    def range(self, lo, hi, count_only=False):
        lo, hi = u32(lo), u32(hi)
        out = np.empty(len(lo), dtype=np.uint32)
        lib().cpu_lsm_range(self._h, _a(lo), _a(hi), len(lo), _a(out),
                            1 if count_only else 0)
        return out

    def rows(self):
        return int(lib().cpu_lsm_rows(self._h))
