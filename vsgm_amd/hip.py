"""ctypes bindings of the device library's C-ABI (include/vsgm_hip.h).

DeviceContext owns a vsgm_ctx (device, stream, memory pool); DeviceArray
is a pool allocation with numpy upload/download. The level operators are
exposed with the vsgm_level_args structure of the header.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from .api import VsgmError, library_paths

MAXQ = 7
_hip = None


def hiplib():
    global _hip
    if _hip is None:
        from .api import lib

        lib()  # loads libvsgm_hip.so globally first
        _hip = C.CDLL(library_paths()[0], mode=C.RTLD_GLOBAL)
        _declare(_hip)
    return _hip


class Cond(C.Structure):
    _fields_ = [("op", C.c_uint32), ("operand", C.c_uint32)]


class LevelArgs(C.Structure):
    _fields_ = [
        ("row_ptrs", C.c_void_p),
        ("cols", C.c_void_p),
        ("vertex_count", C.c_uint64),
        ("path_num", C.c_uint64),
        ("inst", C.c_void_p * MAXQ),
        ("cand", C.c_void_p * MAXQ),
        ("cand_off", C.c_void_p * MAXQ),
        ("cand_idx", C.c_void_p * MAXQ),
        ("cand_idx_base", C.c_uint64 * MAXQ),
        ("target", C.c_uint32),
        ("n_back", C.c_uint32),
        ("back", C.c_uint32 * MAXQ),
        ("n_cond", C.c_uint32),
        ("cond", Cond * MAXQ),
        ("other", C.c_int32),
        ("other_op", C.c_uint32),
        ("hub_slot", C.c_void_p),
        ("hub_bits", C.c_void_p),
        ("hub_words", C.c_uint64),
        ("flags", C.c_uint32),
    ]


class CsrOut(C.Structure):
    _fields_ = [("offsets", C.c_void_p), ("values", C.c_void_p), ("count", C.c_uint64)]


class ExpandOut(C.Structure):
    _fields_ = [("path_num", C.c_uint64), ("inst", C.c_void_p * MAXQ), ("cand_idx", C.c_void_p * MAXQ)]


class OpStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("work", C.c_uint64), ("outputs", C.c_uint64), ("adj_entries", C.c_uint64)]


class ViewOut(C.Structure):
    _fields_ = [("row_ptrs", C.c_void_p), ("cols", C.c_void_p), ("entries", C.c_uint64), ("vertices", C.c_uint64)]


class HubIndex(C.Structure):
    _fields_ = [("slot", C.c_void_p), ("bits", C.c_void_p), ("words", C.c_uint64), ("n_hubs", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("isect_kernel", C.c_uint32), ("max_grid", C.c_uint32), ("cold_weight", C.c_uint32), ("ticket_chunks", C.c_uint32)]


def _declare(L):
    P, u64, i = C.c_void_p, C.c_uint64, C.c_int
    sig = {
        "vsgm_ctx_create": ([i, u64, C.POINTER(P)], i),
        "vsgm_ctx_destroy": ([P], i),
        "vsgm_ctx_sync": ([P], i),
        "vsgm_last_error": ([P], C.c_char_p),
        "vsgm_mem_info": ([P, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)], i),
        "vsgm_ctx_set_budget": ([P, u64], i),
        "vsgm_ctx_bind_thread": ([P], i),
        "vsgm_alloc": ([P, u64, C.POINTER(P)], i),
        "vsgm_free": ([P, P], i),
        "vsgm_h2d": ([P, P, P, u64], i),
        "vsgm_d2h": ([P, P, P, u64], i),
        "vsgm_scan_excl_u64": ([P, P, P, u64], i),
        "vsgm_reduce_u64": ([P, P, u64, P], i),
        "vsgm_gather_u32": ([P, P, u64, P, P], i),
        "vsgm_gather_u64": ([P, P, u64, P, P], i),
        "vsgm_iota_u64": ([P, P, u64, u64], i),
        "vsgm_fill_u64": ([P, P, u64, u64], i),
        "vsgm_intersect_write": ([P, C.POINTER(LevelArgs), C.POINTER(CsrOut)], i),
        "vsgm_intersect_count_total": ([P, C.POINTER(LevelArgs), P], i),
        "vsgm_intersect_count_total_noalloc": ([P, C.POINTER(LevelArgs), P], i),
        "vsgm_intersect_count_paths": ([P, C.POINTER(LevelArgs), P], i),
        "vsgm_expand": ([P, C.POINTER(LevelArgs), P, C.c_uint32, P, C.c_uint32, C.POINTER(ExpandOut)], i),
        "vsgm_filter": ([P, C.POINTER(LevelArgs), C.POINTER(CsrOut)], i),
        "vsgm_count_sets": ([P, C.POINTER(LevelArgs), P], i),
        "vsgm_children_count": ([P, C.POINTER(LevelArgs), i, P], i),
        "vsgm_split_batch": ([P, P, u64, u64, u64, u64, P, u64, C.POINTER(u64)], i),
        "vsgm_last_op_stats": ([P, C.POINTER(OpStats)], i),
        "vsgm_hub_index_build": ([P, P, P, u64, P, C.c_uint32, C.POINTER(HubIndex)], i),
        "vsgm_hub_index_free": ([P, C.POINTER(HubIndex)], i),
        "vsgm_view_self_loops": ([P, P, P, u64, C.POINTER(C.c_uint64)], i),
        "vsgm_csr_from_pairs": ([P, P, u64, u64, i, i, C.POINTER(ViewOut)], i),
        "vsgm_set_tuning": ([P, C.POINTER(Tuning)], i),
        "vsgm_get_tuning": ([P, C.POINTER(Tuning)], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


class DeviceArray:
    def __init__(self, ctx: "DeviceContext", ptr: int, n: int, dtype, owned: bool = True):
        self.ctx, self.ptr, self.n, self.dtype, self.owned = ctx, ptr, n, np.dtype(dtype), owned

    def numpy(self) -> np.ndarray:
        out = np.zeros(self.n, self.dtype)
        if self.n:
            self.ctx.check(hiplib().vsgm_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out

    def free(self):
        if self.owned and self.ptr:
            hiplib().vsgm_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ctx.h:
                self.free()
        except Exception:
            pass


class DeviceContext:
    def __init__(self, device: int = 0, mem_budget: int = 0):
        h = C.c_void_p()
        s = hiplib().vsgm_ctx_create(device, mem_budget, C.byref(h))
        if s != 0:
            raise VsgmError(f"vsgm_ctx_create failed with status {s}")
        self.h = h.value

    def close(self):
        if self.h:
            hiplib().vsgm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def check(self, status, what=""):
        if status != 0:
            raise VsgmError(f"{what}: status {status}: {hiplib().vsgm_last_error(self.h).decode()}")

    def empty(self, n: int, dtype) -> DeviceArray:
        dt = np.dtype(dtype)
        p = C.c_void_p()
        self.check(hiplib().vsgm_alloc(self.h, max(n, 1) * dt.itemsize, C.byref(p)), "alloc")
        return DeviceArray(self, p.value, n, dt)

    def array(self, a) -> DeviceArray:
        a = np.ascontiguousarray(a)
        d = self.empty(a.size, a.dtype)
        if a.size:
            self.check(hiplib().vsgm_h2d(self.h, d.ptr, a.ctypes.data, a.nbytes), "h2d")
            self.sync()
        return d

    def adopt(self, ptr: int, n: int, dtype) -> DeviceArray:
        return DeviceArray(self, ptr, n, dtype, owned=True)

    def sync(self):
        self.check(hiplib().vsgm_ctx_sync(self.h), "sync")

    def tuning(self, **knobs) -> dict:
        """Set any of isect_kernel, max_grid, cold_weight, ticket_chunks
        (vsgm_tuning); returns the setting in force."""
        t = Tuning()
        self.check(hiplib().vsgm_get_tuning(self.h, C.byref(t)), "get_tuning")
        for k, v in knobs.items():
            setattr(t, k, v)
        if knobs:
            self.check(hiplib().vsgm_set_tuning(self.h, C.byref(t)), "set_tuning")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def mem_info(self):
        a, b, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.check(hiplib().vsgm_mem_info(self.h, C.byref(a), C.byref(b), C.byref(p)), "mem_info")
        return a.value, b.value, p.value

    # primitives
    def scan(self, d_in: DeviceArray, d_out: Optional[DeviceArray] = None) -> DeviceArray:
        out = d_out or self.empty(d_in.n + 1, np.uint64)
        self.check(hiplib().vsgm_scan_excl_u64(self.h, d_in.ptr, out.ptr, d_in.n), "scan")
        return out

    def reduce(self, d_in: DeviceArray) -> int:
        tot = self.array(np.zeros(1, np.uint64))
        self.check(hiplib().vsgm_reduce_u64(self.h, d_in.ptr, d_in.n, tot.ptr), "reduce")
        return int(tot.numpy()[0])

    def gather(self, idx: DeviceArray, src: DeviceArray) -> DeviceArray:
        out = self.empty(idx.n, src.dtype)
        fn = hiplib().vsgm_gather_u32 if src.dtype == np.uint32 else hiplib().vsgm_gather_u64
        self.check(fn(self.h, idx.ptr, idx.n, src.ptr, out.ptr), "gather")
        return out

    def iota(self, n: int, base: int = 0) -> DeviceArray:
        out = self.empty(n, np.uint64)
        self.check(hiplib().vsgm_iota_u64(self.h, out.ptr, n, base), "iota")
        return out

    def fill(self, n: int, value: int) -> DeviceArray:
        out = self.empty(n, np.uint64)
        self.check(hiplib().vsgm_fill_u64(self.h, out.ptr, n, value), "fill")
        return out

    def split_batch(self, children: DeviceArray, pf: int, cf: int, batch_bytes: int):
        cap = 1 << 16
        ends = np.zeros(cap, np.uint64)
        n = C.c_uint64()
        self.check(hiplib().vsgm_split_batch(self.h, children.ptr, children.n, pf, cf, batch_bytes, ends.ctypes.data, cap,
                                             C.byref(n)), "split")
        return ends[: n.value].copy()

    def last_stats(self) -> OpStats:
        s = OpStats()
        self.check(hiplib().vsgm_last_op_stats(self.h, C.byref(s)), "stats")
        return s

    def hub_index(self, row_ptrs: DeviceArray, cols: DeviceArray, vertex_count: int, hubs) -> HubIndex:
        """Dense membership index of the lists of `hubs` (free with hub_index_free)."""
        h = np.ascontiguousarray(hubs, dtype=np.uint32)
        out = HubIndex()
        self.check(hiplib().vsgm_hub_index_build(self.h, row_ptrs.ptr, cols.ptr, vertex_count, h.ctypes.data if h.size else None,
                                                 h.size, C.byref(out)), "hub index")
        return out

    def hub_index_free(self, index: HubIndex):
        self.check(hiplib().vsgm_hub_index_free(self.h, C.byref(index)), "hub index free")

    def csr_from_pairs(self, pairs: DeviceArray, vertex_count: int, symmetric: bool = True, drop_isolated: bool = False):
        """Device CSR of an edge list (pairs: u32 device array of 2n ids); returns (row_ptrs, cols) device arrays."""
        out = ViewOut()
        self.check(hiplib().vsgm_csr_from_pairs(self.h, pairs.ptr, pairs.n // 2, vertex_count, int(symmetric), int(drop_isolated),
                                                C.byref(out)), "csr_from_pairs")
        return self.adopt(out.row_ptrs, out.vertices + 1, np.uint64), self.adopt(out.cols, out.entries, np.uint32)
