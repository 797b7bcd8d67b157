"""The C-ABI library loads and exports every entry point include/bnx_rt.h
declares (no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "bnx_rt.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bnx_[a-z0-9_]+)\s*\(", txt)))


def barney_api_functions():
    """the bn* functions include/bnx_barney.h declares"""
    txt = open(os.path.join(ROOT, "include", "bnx_barney.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bn[A-Z][A-Za-z0-9]+)\s*\(", txt)))


def test_library_exports_the_barney_api():
    """every bn* entry point of the mirrored barney.h, with C linkage"""
    import barney_amd
    if not os.path.exists(barney_amd.LIB_PATH):
        pytest.skip("HIP library not built here")
    L = C.CDLL(barney_amd.LIB_PATH)
    fns = barney_api_functions()
    for f in ["bnContextCreate", "bnGeometryCreate", "bnSetData", "bnGroupCreate", "bnSetInstances", "bnBuild",
              "bnRender", "bnFrameBufferRead", "bnVolumeSetXF", "bnCommit"]:
        assert f in fns
    missing = [f for f in fns if not hasattr(L, f)]
    assert not missing, missing


def test_header_declares_core_entry_points():
    fns = header_functions()
    for f in ["bnx_device_create", "bnx_group_build", "bnx_trace", "bnx_render", "bnx_frame_read"]:
        assert f in fns


def test_library_exports_every_declared_symbol():
    import barney_amd
    if not os.path.exists(barney_amd.LIB_PATH):
        pytest.skip("HIP library not built here")
    L = C.CDLL(barney_amd.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(L, f)]
    assert not missing, missing
    assert sorted(barney_amd.EXPORTED) == header_functions()
    assert L.bnx_abi_version() == 1


def test_struct_layouts_match_header():
    from barney_amd import abi
    assert C.sizeof(abi.TrianglesDesc) == 88
    assert C.sizeof(abi.SpheresDesc) == 88
    assert C.sizeof(abi.Renderer) == 48
    assert C.sizeof(abi.RenderStats) == 152
    assert abi.HIT_DTYPE.itemsize == 32 and abi.NODE_DTYPE.itemsize == 64


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    import barney_amd
    monkeypatch.setattr(barney_amd, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(barney_amd, "_lib", None)
    with pytest.raises(RuntimeError):
        barney_amd.lib()


def test_cpp_rtc_mirror_compiles():
    """include/bnx_rtc.hpp (the C++ mirror of barney's rtc interface) is
    self-contained C++17 over bnx_rt.h"""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-I", os.path.join(root, "include"),
                        "-x", "c++", os.path.join(root, "include", "bnx_rtc.hpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_bench_reads_profile_traffic_of_every_workload():
    # the roofline's "traffic" comes from the committed rocprof summaries;
    # trace kernel names hold spaces ("k_trace_static<16, false, false>")
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    for w, k in [("c1", "k_trace_static<32, true, false>"), ("c2", "k_volume_march<0>"),
                 ("c3", "k_trace_static<16, false, false>"), ("c4", "k_volume_march<1>"),
                 ("c5", "k_trace_static<16, false, false>")]:
        t = bench.profile_traffic(w, k)
        assert t is not None and t > 0, (w, k)
    assert bench.profile_traffic("c3", "k_no_such_kernel") is None
