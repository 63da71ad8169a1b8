"""MI355X-native VSGM enumeration path (pm_pipeline) — Python bindings.

The product is native: ``lib/libvsgm_hip.so`` (gfx950 kernels + device
C-ABI, ``include/vsgm_hip.h``) and ``lib/libvsgm.so`` (C++ host: graph IO,
planner, view bins, pipeline executor; C-ABI ``include/vsgm.h``). This
package is a thin ctypes layer over those libraries with the names of the
reference's host classes (Graph, Pattern/Clique/Graphlet, Plan,
ViewBinManager, ViewBinHolder, PipelineExecutor).
"""
from .api import (  # noqa: F401
    Graph,
    Query,
    Plan,
    ViewBinManager,
    ViewBinHolder,
    PipelineExecutor,
    LevelStat,
    VsgmError,
    device_count,
    pm_pipeline,
    preprocess_graph,
    view_packing,
    pack_view_bins,
    kmeans_tool,
    reorder_clusters,
    compact_clusters,
    kmeans_levels,
    fennel,
    library_paths,
    OPS,
)
