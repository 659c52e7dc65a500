"""The C-ABI library loads and exports what include/gvom_abi.h declares (CPU only:
no compute call is made; gvom_create is only called with parameters it rejects
before touching the device)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

import gvom_native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gvom_abi.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(gvom_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S):
        args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        decls[m.group(1)] = len(args)
    return decls


def test_header_declares_the_entry_points():
    d = _declared()
    for name in ("gvom_create", "gvom_destroy", "gvom_process_pointcloud", "gvom_process_batch",
                 "gvom_combine_maps", "gvom_debug_voxel_map", "gvom_get_slot", "gvom_last_error"):
        assert name in d


def test_library_exports_every_declared_symbol():
    lib = nat.load()
    for name, nargs in _declared().items():
        assert hasattr(lib, name), f"{name} not exported"
        assert name in nat.SIGNATURES, f"{name} has no ctypes signature"
        assert len(nat.SIGNATURES[name][1]) == nargs, f"{name}: binding arity differs from the header"
    assert set(nat.SIGNATURES) == set(_declared())


def test_params_struct_layout_matches_header():
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "gvom_abi.h"
int main(void) {
    printf("%zu", sizeof(gvom_params));
#define F(f) printf(" %zu", offsetof(gvom_params, f));
    F(xy_resolution) F(z_resolution) F(xy_size) F(z_size) F(buffer_size) F(reserved0) F(min_distance)
    F(positive_obstacle_threshold) F(negative_obstacle_threshold) F(slope_obstacle_threshold) F(robot_height)
    F(robot_radius) F(ground_to_lidar_height) F(xy_eigen_dist) F(z_eigen_dist)
    return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    P = nat.Params
    assert out[0] == ctypes.sizeof(P)
    assert out[1:] == [getattr(P, f).offset for f, _ in P._fields_]


@pytest.mark.parametrize("field,value,msg", [("xy_size", 2, "xy_size"), ("z_size", 300, "z_size"),
                                             ("buffer_size", 0, "buffer_size"), ("xy_resolution", 0.0, "resolution"),
                                             ("xy_eigen_dist", -1, "eigen")])
def test_create_rejects_bad_parameters_before_touching_the_device(field, value, msg):
    lib = nat.load()
    p = nat.Params(0.4, 0.2, 64, 32, 1, 0, 1.0, 0.5, 0.5, 0.3, 2.0, 4.0, 1.0, 1, 1)
    setattr(p, field, value)
    h = ctypes.c_void_p()
    rc = lib.gvom_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert msg in lib.gvom_last_error(None).decode()


def test_python_class_surfaces_errors():
    import gvom
    with pytest.raises(nat.GvomError, match="buffer_size"):
        gvom.Gvom(0.4, 0.2, 64, 32, 0, 1.0, 0.5, 0.5, 0.3, 2.0, 4.0, 1.0, 1, 1)


def test_reference_method_names_exist():
    import gvom
    for name in ("Process_pointcloud", "process_pointcloud", "combine_maps", "make_debug_voxel_map",
                 "make_debug_height_map", "make_debug_inferred_height_map"):
        assert callable(getattr(gvom.Gvom, name))
    import inspect
    sig = inspect.signature(gvom.Gvom.__init__)
    assert list(sig.parameters)[1:] == [
        "xy_resolution", "z_resolution", "xy_size", "z_size", "buffer_size", "min_distance",
        "positive_obstacle_threshold", "negative_obstacle_threshold", "slope_obsacle_threshold", "robot_height",
        "robot_radius", "ground_to_lidar_height", "xy_eigen_dist", "z_eigen_dist"]


def test_synthetic_workloads_are_deterministic():
    import gvom_synth
    a = gvom_synth.workload(1)
    b = gvom_synth.workload(1)
    assert a.steps[0][0].points.tobytes() == b.steps[0][0].points.tobytes()
    s = a.steps[0][0]
    assert s.points.dtype.name == "float32" and s.transform.dtype.name == "float32"
    assert 0.9 * 8192 < s.points.shape[0] < 8192


def test_library_bound_to_the_runtime_it_was_built_for():
    """The test session loads libgvom_hip.so before torch, so the library runs on the ROCm
    HIP runtime it was built against (same major and minor version), not on torch's bundled
    copy, which carries the same SONAME."""
    import gvom_native as nat
    built, running = nat.hip_versions()
    files = nat.hip_runtime_files()
    print(f"built for HIP {built}, running {running}; runtimes mapped: {files}")
    assert built // 100000 == running // 100000
    assert files
