"""k-means vertex partitioning on the GPU (include/vsgm_partition.h) and the
.vbmap file the view bins are loaded from (ViewBinManager)."""
import ctypes as C
from typing import Optional

import numpy as np

from .api import Graph, VsgmError
from .hip import DeviceContext, hiplib

_declared = False
MAX_K = 16


def _lib():
    global _declared
    L = hiplib()
    if not _declared:
        P, u64, i = C.c_void_p, C.c_uint64, C.c_int
        sig = {
            "vsgm_kmeans_assign": ([P, P, P, u64, i, P, P, P, P], i),
            "vsgm_kmeans_update": ([P, P, P, u64, i, P, P, P, P], i),
            "vsgm_kmeans": ([P, P, P, u64, i, i, u64, P, P, P], i),
            "vsgm_kmeans_members": ([P, P, P, u64, P, u64, i, i, u64, P, P, P], i),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        _declared = True
    return L


def first_vertex(seed: int, vertex_count: int) -> int:
    """The seeded first centroid vertex (SplitMix64 of the seed, mod V)."""
    z = (seed + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return (z ^ (z >> 31)) % vertex_count


class DeviceGraph:
    """A graph's CSR resident in a device context (for the partitioner)."""

    def __init__(self, graph: Graph, ctx: Optional[DeviceContext] = None):
        self.ctx = ctx or DeviceContext(0)
        rp, cl = graph.csr(copy=False)
        self.V = graph.vertex_count
        self.rp = self.ctx.array(rp)
        self.cols = self.ctx.array(cl if len(cl) else np.zeros(1, np.uint32))


def kmeans(graph: Graph, k: int, iters: int = 10, seed: int = 1, ctx: Optional[DeviceContext] = None):
    """Cosine k-means over closed 1-hop neighbourhoods; returns
    (assignment int32[V], sizes u64[k], objective per round)."""
    if not 1 <= k <= MAX_K:
        raise VsgmError(f"k must be in [1, {MAX_K}]")
    dg = DeviceGraph(graph, ctx)
    a = np.zeros(dg.V, np.int32)
    sizes = np.zeros(k, np.uint64)
    obj = np.zeros(iters + 1, np.float64)
    dg.ctx.check(_lib().vsgm_kmeans(dg.ctx.h, dg.rp.ptr, dg.cols.ptr, dg.V, k, iters, seed, a.ctypes.data, sizes.ctypes.data,
                                    obj.ctypes.data), "kmeans")
    return a, sizes, obj


def kmeans_members(graph: Graph, members, k: int, iters: int = 10, seed: int = 1, ctx: Optional[DeviceContext] = None):
    """The run restricted to `members` (features over the whole graph);
    returns (assignment int32[len(members)] in member order, sizes, objective)."""
    if not 1 <= k <= MAX_K:
        raise VsgmError(f"k must be in [1, {MAX_K}]")
    dg = DeviceGraph(graph, ctx)
    mem = np.ascontiguousarray(members, dtype=np.uint32)
    dmem = dg.ctx.array(mem)
    a = np.zeros(mem.size, np.int32)
    sizes = np.zeros(k, np.uint64)
    obj = np.zeros(iters + 1, np.float64)
    dg.ctx.check(_lib().vsgm_kmeans_members(dg.ctx.h, dg.rp.ptr, dg.cols.ptr, dg.V, dmem.ptr, mem.size, k, iters, seed, a.ctypes.data,
                                            sizes.ctypes.data, obj.ctypes.data), "kmeans_members")
    return a, sizes, obj


def assign_step(dg: DeviceGraph, cent: np.ndarray, norm: np.ndarray):
    """One assignment pass for given centroids cent [V][K] (testing hook)."""
    K = cent.shape[1]
    dc, dn = dg.ctx.array(np.ascontiguousarray(cent, np.float64).reshape(-1)), dg.ctx.array(np.ascontiguousarray(norm, np.float64))
    da, ds = dg.ctx.empty(dg.V, np.int32), dg.ctx.empty(dg.V, np.float64)
    dg.ctx.check(_lib().vsgm_kmeans_assign(dg.ctx.h, dg.rp.ptr, dg.cols.ptr, dg.V, K, dc.ptr, dn.ptr, da.ptr, ds.ptr), "assign")
    return da.numpy(), ds.numpy()


def update_step(dg: DeviceGraph, assign: np.ndarray, K: int):
    """One centroid update for a given assignment (testing hook): returns cent [V][K], norms, sizes."""
    da = dg.ctx.array(np.ascontiguousarray(assign, np.int32))
    dc, dn, dz = dg.ctx.empty(dg.V * K, np.float64), dg.ctx.empty(K, np.float64), dg.ctx.empty(K, np.uint64)
    dg.ctx.check(_lib().vsgm_kmeans_update(dg.ctx.h, dg.rp.ptr, dg.cols.ptr, dg.V, K, da.ptr, dc.ptr, dn.ptr, dz.ptr), "update")
    return dc.numpy().reshape(dg.V, K), dn.numpy(), dz.numpy()
