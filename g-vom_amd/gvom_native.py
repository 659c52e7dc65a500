"""ctypes binding of libgvom_hip.so (C-ABI declared in include/gvom_abi.h).

The library is the only compute path: loading fails loudly when it has not been
built, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgvom_hip.so")

GVOM_OK = 0
GVOM_NO_DATA = 1
GVOM_F32 = 0
GVOM_F64 = 1
GVOM_NO_TRANSFORM = -1
GRID_NAMES = ("hard_obstacle_map", "soft_obstacle_map", "ground_certainty_map", "negative_obstacle_map",
              "roughness_map")

# Every symbol the header declares, with (restype, argtypes).
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
SIGNATURES = {
    "gvom_create": (ctypes.c_int, [_P, _P]),
    "gvom_destroy": (None, [_P]),
    "gvom_last_error": (ctypes.c_char_p, [_P]),
    "gvom_reserve": (ctypes.c_int, [_P, _I64, _I64]),
    "gvom_process_pointcloud": (ctypes.c_int, [_P, _P, _I64, _I32, _P, _P, _I32]),
    "gvom_process_pointcloud_device": (ctypes.c_int, [_P, _P, _I64, _I32, _P, _P, _I32]),
    "gvom_process_pointcloud_device_noalloc": (ctypes.c_int, [_P, _P, _I64, _I32, _P, _P, _I32]),
    "gvom_process_batch": (ctypes.c_int, [_P, _P, _P, _I32, _I32, _P, _P, _I32, _I32]),
    "gvom_combine_maps": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "gvom_combine_maps_device": (ctypes.c_int, [_P, _P]),
    "gvom_device_maps": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "gvom_sync": (ctypes.c_int, [_P]),
    "gvom_process_pointcloud2": (ctypes.c_int, [_P, _P, _P, _P, _P, _I32, _I32]),
    "gvom_probe_atomics": (ctypes.c_int, [_P, _I32, _I64, _P]),
    "gvom_occupancy_grids": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _P, _I32, _P]),
    "gvom_debug_voxel_map": (ctypes.c_int, [_P, _P, _I64, _P]),
    "gvom_debug_height_map": (ctypes.c_int, [_P, _P]),
    "gvom_debug_inferred_height_map": (ctypes.c_int, [_P, _P]),
    "gvom_get_slot": (ctypes.c_int, [_P, _I32, _P, _I64, _P, _P, _P, _P, _P, _P]),
    "gvom_get_combined": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P, _P, _P, _P, _P]),
    "gvom_get_maps": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gvom_ring_state": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "gvom_set_timing": (ctypes.c_int, [_P, _I32]),
    "gvom_stage_times": (ctypes.c_int, [_P, _P, _I32]),
    "gvom_stream": (ctypes.c_void_p, [_P]),
    "gvom_set_overlap": (ctypes.c_int, [_P, _I32]),
    "gvom_expand_binary": (ctypes.c_int, [_P, _P, _P, _I32, ctypes.c_double]),
    "gvom_lowpass_binary": (ctypes.c_int, [_P, _P, _P, _I32, _I32, ctypes.c_double]),
    "gvom_convolve": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32]),
    "gvom_diff_forward": (ctypes.c_int, [_P, _P, _I64, _P, _P, ctypes.c_double, _P]),
    "gvom_graph_begin": (ctypes.c_int, [_P, _I32]),
    "gvom_graph_end": (ctypes.c_int, [_P, _P]),
    "gvom_graph_scans": (ctypes.c_int, [_P]),
    "gvom_graph_launch": (ctypes.c_int, [_P, _P, _P, _P, _I32]),
    "gvom_graph_destroy": (None, [_P]),
    "gvom_autotune": (ctypes.c_int, [_P, _P, _I64, _I32, _P, _P, _I32, _P, _P]),
    "gvom_hip_versions": (ctypes.c_int, [_P, _P]),
    "gvom_set_diff_geometry_only": (ctypes.c_int, [_P, _I32]),
    "gvom_hip_runtime_path": (ctypes.c_char_p, []),
    "gvom_set_raycast_config": (ctypes.c_int, [_P, _I32, _I32, _I32]),
    "gvom_set_ray_rings": (ctypes.c_int, [_P, _I64, _I32]),
    "gvom_diff_backward": (ctypes.c_int, [_P, ctypes.c_double, _P, _P, _P]),
    "gvom_diff_forward_async": (ctypes.c_int, [_P, _P, _I64, _P, _P, ctypes.c_double, _P, _P]),
    "gvom_diff_backward_async": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
}


class Params(ctypes.Structure):
    """gvom_params (include/gvom_abi.h): Gvom.__init__'s 14 arguments (gvom.py:21-22)."""
    _fields_ = [
        ("xy_resolution", ctypes.c_double), ("z_resolution", ctypes.c_double),
        ("xy_size", ctypes.c_int32), ("z_size", ctypes.c_int32),
        ("buffer_size", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("min_distance", ctypes.c_double),
        ("positive_obstacle_threshold", ctypes.c_double),
        ("negative_obstacle_threshold", ctypes.c_double),
        ("slope_obstacle_threshold", ctypes.c_double),
        ("robot_height", ctypes.c_double), ("robot_radius", ctypes.c_double),
        ("ground_to_lidar_height", ctypes.c_double),
        ("xy_eigen_dist", ctypes.c_int32), ("z_eigen_dist", ctypes.c_int32),
    ]


class CloudLayout(ctypes.Structure):
    """gvom_cloud_layout (include/gvom_abi.h): where x, y, z sit in PointCloud2 data."""
    _fields_ = [
        ("height", ctypes.c_int64), ("width", ctypes.c_int64),
        ("point_step", ctypes.c_int64), ("row_step", ctypes.c_int64),
        ("offset", ctypes.c_int32 * 3), ("datatype", ctypes.c_int32 * 3),
        ("is_bigendian", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class GvomError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GvomError(f"libgvom_hip.so not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def hip_versions():
    """(built, running): the HIP version libgvom_hip.so was built against and that of the
    HIP runtime it is bound to in this process (major * 10**7 + minor * 10**5 + patch)."""
    built, running = ctypes.c_int32(), ctypes.c_int32()
    load().gvom_hip_versions(ctypes.byref(built), ctypes.byref(running))
    return built.value, running.value


def hip_runtime_files():
    """Distinct HIP runtime files (libamdhip64) mapped into this process.  torch loads its
    bundled copy by the name libamdhip64.so from its own directory, this library by the
    SONAME libamdhip64.so.7: whichever is loaded first serves both names' users that bind
    later by SONAME, so importing torch before this library leaves one runtime, and the
    other order leaves two."""
    files = set()
    with open("/proc/self/maps") as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 6 and "libamdhip64" in os.path.basename(parts[-1]):
                files.add(os.path.realpath(parts[-1]))
    return sorted(files)


def check(rc: int, handle=None) -> int:
    if rc < 0:
        msg = load().gvom_last_error(handle)
        raise GvomError(f"gvom call failed ({rc}): {msg.decode() if msg else ''}")
    return rc


def ptr(a):
    """ctypes pointer of a numpy array (or None)."""
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


_hip = None


def _hip_runtime():
    global _hip
    if _hip is None:
        # the runtime file libgvom_hip.so is bound to: a bare "libamdhip64.so" can name
        # another copy once torch has loaded its own under that name
        path = load().gvom_hip_runtime_path().decode()
        if not path:
            raise GvomError("cannot locate the HIP runtime libgvom_hip.so is bound to")
        _hip = ctypes.CDLL(path)
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
    return _hip


class DeviceArray:
    """A device copy of a numpy array (plain hipMalloc/hipMemcpy), for the _device
    entry points and for benchmarks that start with inputs resident in HBM."""

    def __init__(self, host: "np.ndarray"):
        import numpy as np
        hip = _hip_runtime()
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        self.ptr = ctypes.c_void_p()
        rc = hip.hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1))
        if rc != 0:
            raise GvomError(f"hipMalloc failed ({rc})")
        if self.nbytes:
            rc = hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(ctypes.c_void_p), self.nbytes, 1)
            if rc != 0:
                raise GvomError(f"hipMemcpy failed ({rc})")

    def to_host(self):
        """A host copy of the buffer's current contents (synchronising hipMemcpy), shaped and
        typed like the array it was made from."""
        import numpy as np
        out = np.empty_like(self.host)
        if self.nbytes:
            rc = _hip_runtime().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes, 2)
            if rc != 0:
                raise GvomError(f"hipMemcpy failed ({rc})")
        return out

    def free(self):
        if self.ptr and self.ptr.value:
            _hip_runtime().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
