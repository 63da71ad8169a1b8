"""CPU: the C-ABI libraries load and export every function include/*.h declares."""
import ctypes
import os
import re

from conftest import ROOT

HEADERS = {"vsgm_hip.h": "libvsgm_hip.so", "vsgm_mlp.h": "libvsgm_hip.so", "vsgm_partition.h": "libvsgm_hip.so",
           "vsgm.h": "libvsgm.so"}


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(vsgm_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_headers_declare_entry_points():
    assert len(_declared("vsgm_hip.h")) >= 30
    assert len(_declared("vsgm.h")) >= 40


def test_libraries_export_every_declared_symbol(vsgm):
    hip_path, host_path = vsgm.library_paths()
    libs = {"libvsgm_hip.so": ctypes.CDLL(hip_path), "libvsgm.so": ctypes.CDLL(host_path)}
    missing = []
    for header, lib in HEADERS.items():
        for name in _declared(header):
            if not hasattr(libs[lib], name):
                missing.append(f"{lib}:{name}")
    assert not missing, missing


def test_no_device_is_reported_without_a_gpu(vsgm):
    # on a host without a GPU the device count is 0 and the executor refuses
    # to start rather than falling back to anything
    if vsgm.device_count() == 0:
        g = vsgm.Graph.from_edges(3, [(0, 1), (1, 2)])
        p = vsgm.Plan.optimize(vsgm.Query.clique(3))
        try:
            vsgm.PipelineExecutor(0, g.vertex_count, [p])
        except vsgm.VsgmError as e:
            assert "GPU" in str(e) or "device" in str(e)
        else:
            raise AssertionError("executor started without a GPU")


def test_level_args_layout_matches_header():
    # the ctypes mirror of vsgm_level_args must match the C layout
    from vsgm_amd.hip import LevelArgs

    # ... + flags (u32) and the tail padding to 8-byte alignment
    assert ctypes.sizeof(LevelArgs) == 8 * 4 + 8 * 7 * 5 + 4 * 2 + 4 * 7 + 4 + 8 * 7 + 4 + 4 + 8 * 3 + 8
    assert LevelArgs.flags.offset == 440
