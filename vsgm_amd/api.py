"""ctypes bindings of libvsgm.so (include/vsgm.h).

Each class wraps one opaque handle of the C-ABI and mirrors the reference
class of the same role:
  Graph             src/graph/cpu_graph.h, src/graph/graph_io.h
  Query             src/query/pattern.h, clique.h, graphlet.h, query.h
  Plan              src/query/plan.h (Plan::Optimize)
  ViewBinManager    src/view/view_bin_manager.h, view_bin.h
  ViewBinHolder     src/view/view_bin_holder.h
  PipelineExecutor  src/executor/pipeline_executor.h
Errors raise VsgmError with the library's message. There is no Python
fallback: without the native libraries (or without a GPU, for the
executor) the calls fail.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Iterable, List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBDIR = os.path.join(_HERE, "lib")

OPS = {0: "COMPUTE", 1: "COMPUTE_PATH_COUNT", 2: "MATERIALIZE", 3: "FILTER_COMPUTE", 4: "COUNT", 5: "COMPUTE_COUNT"}


class VsgmError(RuntimeError):
    pass


def library_paths():
    return os.path.join(_LIBDIR, "libvsgm_hip.so"), os.path.join(_LIBDIR, "libvsgm.so")


_lib = None


def lib():
    """Load libvsgm.so (and with it libvsgm_hip.so); raise if not built."""
    global _lib
    if _lib is None:
        hip_path, host_path = library_paths()
        for p in (hip_path, host_path):
            if not os.path.exists(p):
                raise VsgmError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                                f"or `make -C vsgm_amd`")
        C.CDLL(hip_path, mode=C.RTLD_GLOBAL)
        _lib = C.CDLL(host_path, mode=C.RTLD_GLOBAL)
        _declare(_lib)
    return _lib


P = C.c_void_p
u64 = C.c_uint64
PP = C.POINTER(C.c_void_p)


class LevelStat(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("paths", C.c_uint64), ("work", C.c_uint64), ("outputs", C.c_uint64),
                ("adj_entries", C.c_uint64), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double), ("max_ms", C.c_double),
                ("algo_bytes", C.c_uint64), ("work", C.c_uint64), ("paths", C.c_uint64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["name"] = self.name.decode()
        return d


def _declare(L):
    sig = {
        "vsgm_error": ([], C.c_char_p),
        "vsgm_device_count": ([C.POINTER(C.c_int)], C.c_int),
        "vsgm_graph_read_bin": ([C.c_char_p, PP], C.c_int),
        "vsgm_graph_write_bin": ([P, C.c_char_p], C.c_int),
        "vsgm_graph_from_csr": ([u64, P, P, PP], C.c_int),
        "vsgm_graph_from_edges": ([u64, P, u64, PP], C.c_int),
        "vsgm_graph_read_snap": ([C.c_char_p, C.c_int, C.c_int, C.c_int, PP], C.c_int),
        "vsgm_graph_drop_isolated": ([P, PP], C.c_int),
        "vsgm_graph_relabel_degree": ([P, C.c_int, P, PP], C.c_int),
        "vsgm_graph_load": ([C.c_char_p, C.c_int, C.c_int, C.c_int, PP], C.c_int),
        "vsgm_graph_rmat": ([C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, u64, C.c_int, PP], C.c_int),
        "vsgm_graph_chung_lu": ([u64, u64, C.c_double, u64, C.c_int, PP], C.c_int),
        "vsgm_graph_info": ([P, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)], C.c_int),
        "vsgm_graph_csr": ([P, PP, PP], C.c_int),
        "vsgm_graph_validate": ([P, C.c_char_p, C.c_size_t], C.c_int),
        "vsgm_graph_free": ([P], None),
        "vsgm_query_preset": ([C.c_int, PP], C.c_int),
        "vsgm_query_clique": ([C.c_int, PP], C.c_int),
        "vsgm_query_from_matrix": ([C.c_int, P, PP], C.c_int),
        "vsgm_query_graphlets": ([C.c_int, PP, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "vsgm_query_info": ([P, C.POINTER(C.c_int), C.POINTER(C.c_int), P, P], C.c_int),
        "vsgm_query_free": ([P], None),
        "vsgm_plan_optimize": ([P, PP], C.c_int),
        "vsgm_plan_from_order": ([P, P, C.c_int, PP], C.c_int),
        "vsgm_plan_info": ([P, C.POINTER(C.c_int), C.POINTER(u64), C.POINTER(C.c_int)], C.c_int),
        "vsgm_plan_op": ([P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "vsgm_plans_shared_prefix": ([P, P, C.POINTER(C.c_int)], C.c_int),
        "vsgm_plan_order": ([P, P, C.c_int], C.c_int),
        "vsgm_plan_serialize": ([P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
        "vsgm_plan_free": ([P], None),
        "vsgm_vbmap_write": ([C.c_char_p, C.c_int, P, u64], C.c_int),
        "vsgm_view_manager_load": ([P, C.c_int, C.c_int, C.c_char_p, PP], C.c_int),
        "vsgm_view_manager_from_assignment": ([P, C.c_int, C.c_int, C.c_int, P, PP], C.c_int),
        "vsgm_view_manager_info": ([P, C.POINTER(C.c_int), C.POINTER(u64)], C.c_int),
        "vsgm_view_manager_reorder": ([P], C.c_int),
        "vsgm_view_manager_split": ([P, C.c_int, C.c_int], C.c_int),
        "vsgm_view_manager_free": ([P], None),
        "vsgm_view_holder_create": ([u64, u64, PP], C.c_int),
        "vsgm_view_materialize": ([P, C.c_int, P, u64], C.c_int),
        "vsgm_view_holder_info": ([P, C.POINTER(C.c_int), C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)], C.c_int),
        "vsgm_view_holder_csr": ([P, PP, PP, PP], C.c_int),
        "vsgm_view_holder_free": ([P], None),
        "vsgm_executor_create": ([C.c_int, u64, P, C.c_int, u64, u64, PP], C.c_int),
        "vsgm_executor_set_option": ([P, C.c_char_p, C.c_double], C.c_int),
        "vsgm_executor_transfer": ([P, P], C.c_int),
        "vsgm_executor_transfer_graph": ([P, P, P, u64], C.c_int),
        "vsgm_executor_match": ([P, C.c_int], C.c_int),
        "vsgm_executor_counts": ([P, P, C.c_int], C.c_int),
        "vsgm_executor_match_ms": ([P, P, C.c_int], C.c_int),
        "vsgm_executor_shared_stats": ([P, C.POINTER(C.c_double), C.POINTER(u64)], C.c_int),
        "vsgm_executor_reset": ([P], C.c_int),
        "vsgm_executor_level_stats": ([P, C.c_int, P, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "vsgm_executor_peak_bytes": ([P, C.POINTER(u64)], C.c_int),
        "vsgm_executor_hub_count": ([P, C.POINTER(C.c_uint32)], C.c_int),
        "vsgm_executor_set_plan_tuning": ([P, C.c_int, C.c_uint32, C.c_uint32, u64, C.c_uint32, C.c_uint32], C.c_int),
        "vsgm_executor_get_plan_tuning": ([P, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(u64),
                                           C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)], C.c_int),
        "vsgm_executor_add_view": ([P, P, C.POINTER(C.c_int)], C.c_int),
        "vsgm_executor_set_graph": ([P, P], C.c_int),
        "vsgm_executor_add_device_view": ([P, P, u64, C.c_int, u64, C.POINTER(C.c_int)], C.c_int),
        "vsgm_executor_transfer_device": ([P, P, u64, C.c_int, u64], C.c_int),
        "vsgm_executor_prefetch_device": ([P, P, u64, C.c_int, u64], C.c_int),
        "vsgm_executor_view_csr": ([P, C.c_int, P, P, u64, C.POINTER(u64)], C.c_int),
        "vsgm_executor_match_all": ([P, P, C.c_int], C.c_int),
        "vsgm_executor_capture": ([P, P, C.c_int, C.POINTER(u64)], C.c_int),
        "vsgm_executor_replay": ([P, C.POINTER(C.c_double)], C.c_int),
        "vsgm_executor_context": ([P, PP], C.c_int),
        "vsgm_profile_enable": ([P, C.c_int], C.c_int),
        "vsgm_profile_collect": ([P, P, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "vsgm_last_error": ([P], C.c_char_p),
        "vsgm_executor_free": ([P], None),
        "vsgm_pm_pipeline": ([C.c_int, C.POINTER(C.c_char_p)], C.c_int),
        "vsgm_preprocess_graph": ([C.c_int, C.POINTER(C.c_char_p)], C.c_int),
        "vsgm_view_packing": ([C.c_int, C.POINTER(C.c_char_p)], C.c_int),
        "vsgm_kmeans_tool": ([C.c_int, C.POINTER(C.c_char_p)], C.c_int),
        "vsgm_clusters_reorder": ([P, C.c_int, P], C.c_int),
        "vsgm_graph_fennel": ([P, C.c_int, C.c_int, P], C.c_int),
        "vsgm_graph_kmeans": ([P, C.c_char_p, C.c_int, u64, C.c_int, C.c_int, P, C.POINTER(C.c_int)], C.c_int),
        "vsgm_clusters_compact": ([C.c_int, P, u64, C.POINTER(C.c_int)], C.c_int),
        "vsgm_view_pack": ([P, C.c_int, P, C.c_int, C.c_int, u64, C.c_int, C.c_int, C.c_int, P, C.POINTER(C.c_int), P, P, C.c_int],
                           C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


def _check(status: int, what: str = ""):
    if status != 0:
        msg = lib().vsgm_error().decode()
        raise VsgmError(f"{what}: status {status}: {msg}")


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().vsgm_device_count(C.byref(n)), "device_count")
    return n.value


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class _Handle:
    _free = None

    def __init__(self, h):
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and self._free:
            getattr(_lib, self._free)(h)
            self._h = None

    @property
    def handle(self):
        return self._h


class Graph(_Handle):
    """Undirected CSR data graph (src/graph/cpu_graph.h)."""

    _free = "vsgm_graph_free"

    @classmethod
    def read_bin(cls, path: str) -> "Graph":
        h = C.c_void_p()
        _check(lib().vsgm_graph_read_bin(path.encode(), C.byref(h)), f"read {path}")
        return cls(h.value)

    def write_bin(self, path: str):
        _check(lib().vsgm_graph_write_bin(self._h, path.encode()), f"write {path}")

    @classmethod
    def from_csr(cls, row_ptrs, cols) -> "Graph":
        rp = np.ascontiguousarray(row_ptrs, dtype=np.uint64)
        cl = np.ascontiguousarray(cols, dtype=np.uint32)
        h = C.c_void_p()
        _check(lib().vsgm_graph_from_csr(len(rp) - 1, _ptr(rp), _ptr(cl), C.byref(h)), "from_csr")
        return cls(h.value)

    @classmethod
    def from_edges(cls, vertex_count: int, pairs) -> "Graph":
        pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1, 2))
        h = C.c_void_p()
        _check(lib().vsgm_graph_from_edges(vertex_count, _ptr(pr), pr.shape[0], C.byref(h)), "from_edges")
        return cls(h.value)

    @classmethod
    def read_snap(cls, path: str, directed: bool = False, device: int = -1, threads: int = 0) -> "Graph":
        """SNAP edge list -> cleaned graph without isolated vertices; device >= 0 builds the CSR on that GPU."""
        h = C.c_void_p()
        _check(lib().vsgm_graph_read_snap(path.encode(), int(directed), device, threads, C.byref(h)), f"read {path}")
        return cls(h.value)

    @classmethod
    def load(cls, path: str, device: int = -1, threads: int = 0, drop_isolated: bool = False) -> "Graph":
        """Graph file by suffix: .bin CSR, .txt SNAP, .lvid large-id SNAP."""
        h = C.c_void_p()
        _check(lib().vsgm_graph_load(path.encode(), device, threads, int(drop_isolated), C.byref(h)), f"load {path}")
        return cls(h.value)

    def drop_isolated(self) -> "Graph":
        h = C.c_void_p()
        _check(lib().vsgm_graph_drop_isolated(self._h, C.byref(h)), "drop_isolated")
        return Graph(h.value)

    def relabel_by_degree(self, threads: int = 0):
        """(isomorphic copy with ids in ascending (degree, id) order, new_id u32[V]): vertex v is new_id[v] in the
        copy; pattern counts are unchanged."""
        ids = np.empty(self.vertex_count, np.uint32)
        h = C.c_void_p()
        _check(lib().vsgm_graph_relabel_degree(self._h, threads, _ptr(ids), C.byref(h)), "relabel_degree")
        return Graph(h.value), ids

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1, threads: int = 0) -> "Graph":
        h = C.c_void_p()
        _check(lib().vsgm_graph_rmat(scale, edge_factor, a, b, c, seed, threads, C.byref(h)), "rmat")
        return cls(h.value)

    @classmethod
    def chung_lu(cls, vertex_count: int, draws: int, gamma: float = 2.3, seed: int = 1, threads: int = 0) -> "Graph":
        h = C.c_void_p()
        _check(lib().vsgm_graph_chung_lu(vertex_count, draws, gamma, seed, threads, C.byref(h)), "chung_lu")
        return cls(h.value)

    def _info(self):
        v, e, m = u64(), u64(), u64()
        _check(lib().vsgm_graph_info(self._h, C.byref(v), C.byref(e), C.byref(m)), "info")
        return v.value, e.value, m.value

    @property
    def vertex_count(self) -> int:
        return self._info()[0]

    @property
    def entries(self) -> int:
        return self._info()[1]

    @property
    def edge_count(self) -> int:
        return self._info()[1] // 2

    @property
    def max_degree(self) -> int:
        return self._info()[2]

    def csr(self, copy: bool = True):
        """(row_ptrs u64[V+1], cols u32[entries]) as numpy arrays."""
        v, e, _ = self._info()
        rp, cl = C.c_void_p(), C.c_void_p()
        _check(lib().vsgm_graph_csr(self._h, C.byref(rp), C.byref(cl)), "csr")
        r = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_uint64)), shape=(v + 1,))
        c = np.ctypeslib.as_array(C.cast(cl, C.POINTER(C.c_uint32)), shape=(max(e, 1),))[:e] if e else np.zeros(0, np.uint32)
        return (r.copy(), c.copy()) if copy else (r, c)

    def validate(self) -> str:
        buf = C.create_string_buffer(256)
        _check(lib().vsgm_graph_validate(self._h, buf, 256), "validate")
        return buf.value.decode()


class Query(_Handle):
    """A pattern query: preset (Pattern), Clique, Graphlet, or any connected shape."""

    _free = "vsgm_query_free"

    @classmethod
    def preset(cls, k: int) -> "Query":
        h = C.c_void_p()
        _check(lib().vsgm_query_preset(k, C.byref(h)), f"preset P{k}")
        return cls(h.value)

    @classmethod
    def clique(cls, k: int) -> "Query":
        h = C.c_void_p()
        _check(lib().vsgm_query_clique(k, C.byref(h)), f"clique {k}")
        return cls(h.value)

    @classmethod
    def from_matrix(cls, adj) -> "Query":
        a = np.ascontiguousarray(np.asarray(adj, dtype=np.uint8))
        n = a.shape[0]
        h = C.c_void_p()
        _check(lib().vsgm_query_from_matrix(n, _ptr(a), C.byref(h)), "from_matrix")
        return cls(h.value)

    @classmethod
    def graphlets(cls, k: int) -> List["Query"]:
        n = C.c_int(0)
        arr = (C.c_void_p * 64)()
        _check(lib().vsgm_query_graphlets(k, arr, 64, C.byref(n)), f"graphlets {k}")
        return [cls(arr[i]) for i in range(n.value)]

    def info(self):
        n, m = C.c_int(), C.c_int()
        _check(lib().vsgm_query_info(self._h, C.byref(n), C.byref(m), None, None), "query info")
        adj = np.zeros((n.value, n.value), np.uint8)
        cond = np.zeros((n.value, n.value), np.uint8)
        _check(lib().vsgm_query_info(self._h, None, None, _ptr(adj), _ptr(cond)), "query info")
        return n.value, m.value, adj, cond

    @property
    def vertex_count(self) -> int:
        return self.info()[0]

    @property
    def adjacency(self) -> np.ndarray:
        return self.info()[2]

    @property
    def conditions(self) -> np.ndarray:
        return self.info()[3]


class Plan(_Handle):
    """Execution plan (src/query/plan.h). Plan.optimize(q) is Plan(q).Optimize()."""

    _free = "vsgm_plan_free"

    def __init__(self, h, query: Query):
        super().__init__(h)
        self.query = query  # keep the query alive

    @classmethod
    def optimize(cls, query: Query) -> "Plan":
        h = C.c_void_p()
        _check(lib().vsgm_plan_optimize(query.handle, C.byref(h)), "optimize")
        return cls(h.value, query)

    @classmethod
    def from_order(cls, query: Query, order: Sequence[int]) -> "Plan":
        s = np.ascontiguousarray(np.asarray(order, dtype=np.uint32))
        h = C.c_void_p()
        _check(lib().vsgm_plan_from_order(query.handle, _ptr(s), len(s), C.byref(h)), f"from_order {list(order)}")
        return cls(h.value, query)

    def _info(self):
        hop, rd, n = C.c_int(), u64(), C.c_int()
        _check(lib().vsgm_plan_info(self._h, C.byref(hop), C.byref(rd), C.byref(n)), "plan info")
        return hop.value, rd.value, n.value

    @property
    def hop(self) -> int:
        return self._info()[0]

    @property
    def root_degree(self) -> int:
        return self._info()[1]

    def shared_prefix(self, other: "Plan") -> int:
        """Leading exec levels this plan shares with `other` (the unit of prefix-shared matching)."""
        n = C.c_int()
        _check(lib().vsgm_plans_shared_prefix(self._h, other._h, C.byref(n)), "shared prefix")
        return n.value

    @property
    def ops(self):
        out = []
        for i in range(self._info()[2]):
            op, v = C.c_int(), C.c_int()
            _check(lib().vsgm_plan_op(self._h, i, C.byref(op), C.byref(v)), "plan op")
            out.append((OPS[op.value], v.value))
        return out

    @property
    def order(self) -> List[int]:
        n = self.query.vertex_count
        s = np.zeros(n, np.uint32)
        _check(lib().vsgm_plan_order(self._h, _ptr(s), n), "plan order")
        return [int(x) for x in s]

    def serialize(self) -> str:
        need = C.c_size_t()
        _check(lib().vsgm_plan_serialize(self._h, None, 0, C.byref(need)), "serialize")
        buf = C.create_string_buffer(need.value)
        _check(lib().vsgm_plan_serialize(self._h, buf, need.value, None), "serialize")
        return buf.value.decode()


class ViewBinHolder(_Handle):
    """Pinned buffers of one materialised view bin (src/view/view_bin_holder.h)."""

    _free = "vsgm_view_holder_free"

    def __init__(self, vertex_count: int, max_bytes: int):
        h = C.c_void_p()
        _check(lib().vsgm_view_holder_create(vertex_count, max_bytes, C.byref(h)), "holder")
        super().__init__(h.value)

    def info(self):
        b, s, e, v = C.c_int(), u64(), u64(), u64()
        _check(lib().vsgm_view_holder_info(self._h, C.byref(b), C.byref(s), C.byref(e), C.byref(v)), "holder info")
        return {"bin": b.value, "sources": s.value, "entries": e.value, "vertices": v.value}

    def csr(self, vertex_count: int):
        i = self.info()
        rp, cl, sr = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib().vsgm_view_holder_csr(self._h, C.byref(rp), C.byref(cl), C.byref(sr)), "holder csr")
        r = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_uint64)), shape=(vertex_count + 1,)).copy()
        c = (np.ctypeslib.as_array(C.cast(cl, C.POINTER(C.c_uint32)), shape=(i["entries"],)).copy()
             if i["entries"] else np.zeros(0, np.uint32))
        s = (np.ctypeslib.as_array(C.cast(sr, C.POINTER(C.c_uint32)), shape=(i["sources"],)).copy()
             if i["sources"] else np.zeros(0, np.uint32))
        return r, c, s


class ViewBinManager(_Handle):
    """View-bin partition and materialisation (src/view/view_bin_manager.h)."""

    _free = "vsgm_view_manager_free"

    def __init__(self, h, graph: Graph):
        super().__init__(h)
        self.graph = graph

    @classmethod
    def load(cls, graph: Graph, hop: int, vbmap: str, threads: int = 0) -> "ViewBinManager":
        h = C.c_void_p()
        _check(lib().vsgm_view_manager_load(graph.handle, hop, threads, vbmap.encode(), C.byref(h)), f"load {vbmap}")
        return cls(h.value, graph)

    @classmethod
    def from_assignment(cls, graph: Graph, hop: int, n_bins: int, assignment, threads: int = 0) -> "ViewBinManager":
        a = np.ascontiguousarray(np.asarray(assignment, dtype=np.int32))
        h = C.c_void_p()
        _check(lib().vsgm_view_manager_from_assignment(graph.handle, hop, threads, n_bins, _ptr(a), C.byref(h)), "partition")
        return cls(h.value, graph)

    @staticmethod
    def write_vbmap(path: str, n_bins: int, assignment):
        a = np.ascontiguousarray(np.asarray(assignment, dtype=np.int32))
        _check(lib().vsgm_vbmap_write(path.encode(), n_bins, _ptr(a), len(a)), f"write {path}")

    def _info(self):
        n, m = C.c_int(), u64()
        _check(lib().vsgm_view_manager_info(self._h, C.byref(n), C.byref(m)), "manager info")
        return n.value, m.value

    @property
    def n_bins(self) -> int:
        return self._info()[0]

    @property
    def max_sources(self) -> int:
        return self._info()[1]

    def reorder(self):
        _check(lib().vsgm_view_manager_reorder(self._h), "reorder")

    def split(self, devices: int, times: int = 1):
        _check(lib().vsgm_view_manager_split(self._h, devices, times), "split")

    def materialize(self, index: int, holder: ViewBinHolder, root_degree: int = 0):
        _check(lib().vsgm_view_materialize(self._h, index, holder.handle, root_degree), f"materialize bin {index}")


class PipelineExecutor(_Handle):
    """One device's enumeration engine (src/executor/pipeline_executor.h)."""

    _free = "vsgm_executor_free"

    def __init__(self, device: int, vertex_count: int, plans: Sequence[Plan], max_sources: int = 0, mem_budget: int = 0):
        self.plans = list(plans)
        arr = (C.c_void_p * len(self.plans))(*[p.handle for p in self.plans])
        h = C.c_void_p()
        _check(lib().vsgm_executor_create(device, vertex_count, arr, len(self.plans), max_sources, mem_budget, C.byref(h)),
               "executor")
        super().__init__(h.value)

    def set_option(self, key: str, value: float):
        _check(lib().vsgm_executor_set_option(self._h, key.encode(), float(value)), f"option {key}")

    def transfer(self, holder: ViewBinHolder):
        _check(lib().vsgm_executor_transfer(self._h, holder.handle), "transfer")

    def transfer_graph(self, graph: Graph, sources: Optional[Iterable[int]] = None):
        if sources is None:
            _check(lib().vsgm_executor_transfer_graph(self._h, graph.handle, None, 0), "transfer_graph")
        else:
            s = np.ascontiguousarray(np.asarray(list(sources) if not isinstance(sources, np.ndarray) else sources,
                                                dtype=np.uint32))
            _check(lib().vsgm_executor_transfer_graph(self._h, graph.handle, _ptr(s), len(s)), "transfer_graph")

    def match(self, do_match: bool = True):
        _check(lib().vsgm_executor_match(self._h, int(do_match)), "match")

    def reset(self):
        _check(lib().vsgm_executor_reset(self._h), "reset")

    @property
    def total_counts(self) -> List[int]:
        out = np.zeros(len(self.plans), np.uint64)
        _check(lib().vsgm_executor_counts(self._h, _ptr(out), len(self.plans)), "counts")
        return [int(x) for x in out]

    @property
    def match_ms(self) -> List[float]:
        out = np.zeros(len(self.plans), np.float64)
        _check(lib().vsgm_executor_match_ms(self._h, _ptr(out), len(self.plans)), "match_ms")
        return [float(x) for x in out]

    def shared_stats(self) -> dict:
        """Prefix-shared matching: summed time of the shared calls and the level-op runs saved."""
        ms, runs = C.c_double(), u64()
        _check(lib().vsgm_executor_shared_stats(self._h, C.byref(ms), C.byref(runs)), "shared_stats")
        return {"ms": ms.value, "saved_runs": runs.value}

    def level_stats(self, plan_index: int) -> List[dict]:
        arr = (LevelStat * 64)()
        n = C.c_int()
        _check(lib().vsgm_executor_level_stats(self._h, plan_index, arr, 64, C.byref(n)), "level stats")
        return [arr[i].as_dict() for i in range(n.value)]

    @property
    def peak_bytes(self) -> int:
        b = u64()
        _check(lib().vsgm_executor_peak_bytes(self._h, C.byref(b)), "peak")
        return b.value

    def set_plan_tuning(self, plan: int, isect_kernel: int = 4, max_grid: int = 8192, batch_bytes: int = 0, cold_weight: int = 16,
                        ticket_chunks: int = 64):
        _check(lib().vsgm_executor_set_plan_tuning(self._h, plan, isect_kernel, max_grid, batch_bytes, cold_weight, ticket_chunks),
               "plan tuning")

    def plan_tuning(self, plan: int) -> dict:
        k, g, b, w, t = C.c_uint32(), C.c_uint32(), u64(), C.c_uint32(), C.c_uint32()
        _check(lib().vsgm_executor_get_plan_tuning(self._h, plan, C.byref(k), C.byref(g), C.byref(b), C.byref(w), C.byref(t)),
               "plan tuning")
        return {"isect_kernel": k.value, "max_grid": g.value, "batch_bytes": b.value, "cold_weight": w.value, "ticket_chunks": t.value}

    def add_view(self, holder: ViewBinHolder) -> int:
        """Keep one more view bin resident (batched mode)."""
        i = C.c_int()
        _check(lib().vsgm_executor_add_view(self._h, holder.handle, C.byref(i)), "add_view")
        return i.value

    def set_graph(self, graph: Graph):
        """Keep the whole graph resident for device-built views."""
        _check(lib().vsgm_executor_set_graph(self._h, graph.handle), "set_graph")

    def add_device_view(self, sources, hop: int, root_degree: int) -> int:
        """Build one bin's view on the device (from the set_graph graph) and keep it resident."""
        s = np.ascontiguousarray(np.asarray(sources, dtype=np.uint32))
        i = C.c_int()
        _check(lib().vsgm_executor_add_device_view(self._h, _ptr(s), len(s), hop, root_degree, C.byref(i)), "add_device_view")
        return i.value

    def transfer_device(self, sources, hop: int, root_degree: int):
        """Build one bin's view on the device and make it the only resident view (like transfer)."""
        s = np.ascontiguousarray(np.asarray(sources, dtype=np.uint32))
        _check(lib().vsgm_executor_transfer_device(self._h, _ptr(s), len(s), hop, root_degree), "transfer_device")

    def prefetch_device(self, sources, hop: int, root_degree: int):
        """Start building a bin's view in the background (second device context and stream); the next
        transfer_device with the same arguments takes it, so the build overlaps the current match."""
        s = np.ascontiguousarray(np.asarray(sources, dtype=np.uint32))
        _check(lib().vsgm_executor_prefetch_device(self._h, _ptr(s), len(s), hop, root_degree), "prefetch_device")

    def view_csr(self, view: int, vertex_count: int):
        n = u64()
        rp = np.zeros(vertex_count + 1, np.uint64)
        _check(lib().vsgm_executor_view_csr(self._h, view, _ptr(rp), None, 0, C.byref(n)), "view_csr")
        cl = np.zeros(max(n.value, 1), np.uint32)
        _check(lib().vsgm_executor_view_csr(self._h, view, None, _ptr(cl), cl.size, C.byref(n)), "view_csr")
        return rp, cl[: n.value]

    @staticmethod
    def _plan_list(plans):
        if plans is None:
            return None, 0
        a = np.ascontiguousarray(np.asarray(list(plans), dtype=np.int32))
        return a, len(a)

    def match_all(self, plans: Optional[Sequence[int]] = None):
        """The listed plans (all when None) over every resident view."""
        a, n = self._plan_list(plans)
        _check(lib().vsgm_executor_match_all(self._h, _ptr(a) if a is not None else None, n), "match_all")

    def capture(self, plans: Optional[Sequence[int]] = None) -> int:
        """Record one match_all, then capture it into one device graph;
        returns the graph's node count (the recorded run's counts are added)."""
        a, n = self._plan_list(plans)
        nodes = u64()
        _check(lib().vsgm_executor_capture(self._h, _ptr(a) if a is not None else None, n, C.byref(nodes)), "capture")
        return nodes.value

    def replay(self) -> float:
        """Launch the captured graph once; returns its device time in ms."""
        ms = C.c_double()
        _check(lib().vsgm_executor_replay(self._h, C.byref(ms)), "replay")
        return ms.value

    def hub_count(self) -> int:
        """Vertices of the current view in the dense membership index."""
        n = C.c_uint32()
        _check(lib().vsgm_executor_hub_count(self._h, C.byref(n)), "hub count")
        return n.value

    def _ctx(self):
        c = C.c_void_p()
        _check(lib().vsgm_executor_context(self._h, C.byref(c)), "context")
        return c.value

    def profile(self, enable: bool = True):
        """Bracket every level-op kernel launch with HIP events (on the executor's stream)."""
        ctx = self._ctx()
        if lib().vsgm_profile_enable(ctx, int(enable)) != 0:
            raise VsgmError(lib().vsgm_last_error(ctx).decode())

    def kernel_stats(self) -> List[dict]:
        ctx = self._ctx()
        arr = (KernelStat * 64)()
        n = C.c_int()
        if lib().vsgm_profile_collect(ctx, arr, 64, C.byref(n)) != 0:
            raise VsgmError(lib().vsgm_last_error(ctx).decode())
        return [arr[i].as_dict() for i in range(min(n.value, 64))]


def pm_pipeline(argv: Sequence[str]) -> int:
    """Run the pm_pipeline driver in-process with reference-style flags."""
    args = ["pm_pipeline"] + list(argv)
    arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
    return lib().vsgm_pm_pipeline(len(args), arr)


def preprocess_graph(argv: Sequence[str]) -> int:
    """Run the preprocess_graph tool in-process with reference-style flags."""
    args = ["preprocess_graph"] + list(argv)
    arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
    return lib().vsgm_preprocess_graph(len(args), arr)


def view_packing(argv: Sequence[str]) -> int:
    """Run the view_packing tool in-process with reference-style flags."""
    args = ["view_packing"] + list(argv)
    arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
    return lib().vsgm_view_packing(len(args), arr)


def pack_view_bins(graph: "Graph", hop: int, partition, first_layer: int, second_layer: int = 0, max_bin_entries: int = 2500000000,
                   pool_size: int = 1, degree_order: bool = False, threads: int = 0):
    """Pack a vertex partition into view bins; returns (bins, assignment int32[V], sources u64[bins], sizes u64[bins])."""
    part = np.ascontiguousarray(partition, dtype=np.int32)
    V = graph.vertex_count
    if part.shape != (V,):
        raise VsgmError("partition must hold one id per vertex")
    assign = np.zeros(V, np.int32)
    bins = C.c_int()
    cap = V + 1
    src, size = np.zeros(cap, np.uint64), np.zeros(cap, np.uint64)
    _check(lib().vsgm_view_pack(graph._h, hop, _ptr(part), first_layer, second_layer, max_bin_entries, pool_size, int(degree_order),
                                threads, _ptr(assign), C.byref(bins), _ptr(src), _ptr(size), cap), "view_pack")
    n = bins.value
    return n, assign, src[:n].copy(), size[:n].copy()


def kmeans_tool(argv: Sequence[str]) -> int:
    """Run the kmeans tool (GPU) in-process with reference-style flags."""
    args = ["kmeans"] + list(argv)
    arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
    return lib().vsgm_kmeans_tool(len(args), arr)


def reorder_clusters(graph: "Graph", k: int, assignment) -> np.ndarray:
    a = np.array(assignment, dtype=np.int32)
    _check(lib().vsgm_clusters_reorder(graph._h, k, _ptr(a)), "clusters_reorder")
    return a


def compact_clusters(k: int, assignment):
    a = np.array(assignment, dtype=np.int32)
    kept = C.c_int()
    _check(lib().vsgm_clusters_compact(k, _ptr(a), a.size, C.byref(kept)), "clusters_compact")
    return kept.value, a


def kmeans_levels(graph: "Graph", levels: str, iters: int = 10, seed: int = 1, device: int = 0, reorder: bool = False):
    """Multi-level GPU k-means ("K", "AxB...", "A^X"); returns (assignment int32[V], id range)."""
    a = np.zeros(graph.vertex_count, np.int32)
    n = C.c_int()
    _check(lib().vsgm_graph_kmeans(graph._h, levels.encode(), iters, seed, device, int(reorder), _ptr(a), C.byref(n)), "graph_kmeans")
    return a, n.value


def fennel(graph: "Graph", parts: int, hop: int = 1) -> np.ndarray:
    """Streaming Fennel partition (the reference's fennel tool); int32[V]."""
    a = np.zeros(graph.vertex_count, np.int32)
    _check(lib().vsgm_graph_fennel(graph._h, parts, hop, _ptr(a)), "graph_fennel")
    return a
