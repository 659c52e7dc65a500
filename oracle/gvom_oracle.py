"""CPU oracle for G-VOM's scan->map path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker.  The product (g-vom_amd/) never imports it.

`OracleGvom` mirrors the reference class `Gvom` (scripts/gvom.py:8-442): the same
14 constructor arguments, `Process_pointcloud` / `combine_maps`, and the same ring
buffer bookkeeping (gvom.py:198-216, 225-308).  The per-voxel arithmetic runs in
liboracle.so (gvom_oracle.c, exact Numba-CUDA typing, no FMA contraction).

Every slot and the combined map are kept in the reference's own data format:
an int32[V] index map in reference layout (v = x + y*xy + z*xy*xy) and encoding
(>= 0 cell id, -1 unknown, -t-1 free with t ray passes), plus compact per-cell
arrays.  `canonical_*` helpers turn that into per-voxel dense fields so that two
implementations with different (equally legal) cell-id orders can be compared.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Params(ctypes.Structure):
    _fields_ = [
        ("xy_resolution", ctypes.c_double), ("z_resolution", ctypes.c_double),
        ("xy_size", ctypes.c_int32), ("z_size", ctypes.c_int32),
        ("buffer_size", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("min_distance", ctypes.c_double),
        ("positive_obstacle_threshold", ctypes.c_double),
        ("negative_obstacle_threshold", ctypes.c_double),
        ("slope_obstacle_threshold", ctypes.c_double),
        ("robot_height", ctypes.c_double), ("robot_radius", ctypes.c_double),
        ("ground_to_lidar_height", ctypes.c_double),
        ("xy_eigen_dist", ctypes.c_int32), ("z_eigen_dist", ctypes.c_int32),
    ]


class _Source(ctypes.Structure):
    _fields_ = [
        ("index_map", ctypes.c_void_p), ("hit", ctypes.c_void_p), ("total", ctypes.c_void_p),
        ("min_h", ctypes.c_void_p), ("metrics", ctypes.c_void_p), ("origin", ctypes.c_void_p),
    ]


def load_library():
    """Load liboracle.so (built by oracle/Makefile)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    lib.orc_process.restype = ctypes.c_int64
    lib.orc_process.argtypes = [P, P, ctypes.c_int64, ctypes.c_int32, P, P, ctypes.c_int32,
                                P, P, P, P, P, P, P, ctypes.c_int64, P, P, P]
    lib.orc_combine.restype = ctypes.c_int64
    lib.orc_combine.argtypes = [P, ctypes.c_int32, P, P, P, P, P, P, P, P, P, P, ctypes.c_int64,
                                P, P, P, P, P, P, P, P, P]
    lib.orc_origin.restype = None
    lib.orc_origin.argtypes = [P, P, P]
    lib.orc_eigenvalues.restype = None
    lib.orc_eigenvalues.argtypes = [P, P, ctypes.c_int64]
    lib.orc_slope_guess.restype = None
    lib.orc_slope_guess.argtypes = [P, P, P, P, P, P, P]
    lib.orc_set_planted_mistake.restype = None
    lib.orc_set_planted_mistake.argtypes = [ctypes.c_int]
    _LIB = lib
    return lib


# Planted mistakes (pinning tests only): a pin is shown to fail when the oracle makes the
# mistake it guards against.  1-5 live in the C restatement (see gvom_oracle.c), 6-7 in
# the debug exporters below: 6 = hit ratio inverted (total / hit), 7 = planarity column
# taken as lambda0 - lambda2 (gvom.py:492-499), 8 = height-map rows indexed x * xy + y
# instead of x + y * xy (gvom.py:460, 475).
_PLANTED = 0


class planted_mistake:
    """Context manager: run the oracle with planted mistake m (0 = none)."""

    def __init__(self, m):
        self.m = int(m)

    def __enter__(self):
        global _PLANTED
        _PLANTED = self.m
        load_library().orc_set_planted_mistake(self.m)
        return self

    def __exit__(self, *exc):
        global _PLANTED
        _PLANTED = 0
        load_library().orc_set_planted_mistake(0)
        return False


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class ScanSlot:
    """One ring slot in the reference's format (gvom.py:198-208)."""
    index_map: np.ndarray   # int32[V]
    hit: np.ndarray         # int32[C_s]
    total: np.ndarray       # int32[C_s]
    metrics: np.ndarray     # f64[C_s, 10]
    min_h: np.ndarray       # f32[C_s]
    origin: np.ndarray      # f64[3]
    points_world: np.ndarray
    valid: np.ndarray


@dataclass
class CombinedMap:
    index_map: np.ndarray   # int32[V]
    hit: np.ndarray         # int32[C]
    total: np.ndarray       # int32[C]
    min_h: np.ndarray       # f32[C]
    metrics: np.ndarray     # f32[C, 10]
    eigen: np.ndarray       # f32[C, 3]
    origin: np.ndarray      # f64[3]
    height: np.ndarray      # f64[xy, xy]
    inferred: np.ndarray
    slope_x: np.ndarray
    slope_y: np.ndarray
    rough: np.ndarray
    delta: np.ndarray
    pos: np.ndarray         # int32[xy, xy]
    neg: np.ndarray
    vis: np.ndarray


def _dtype_code(a):
    if a.dtype == np.float32:
        return 0
    if a.dtype == np.float64:
        return 1
    raise TypeError(f"unsupported dtype {a.dtype}")


class OracleGvom:
    """Oracle with the reference class's constructor and methods (gvom.py:21-442)."""

    def __init__(self, xy_resolution, z_resolution, xy_size, z_size, buffer_size, min_distance,
                 positive_obstacle_threshold, negative_obstacle_threshold, slope_obsacle_threshold,
                 robot_height, robot_radius, ground_to_lidar_height, xy_eigen_dist, z_eigen_dist):
        self.lib = load_library()
        self.p = _Params(float(xy_resolution), float(z_resolution), int(xy_size), int(z_size),
                         int(buffer_size), 0, float(min_distance), float(positive_obstacle_threshold),
                         float(negative_obstacle_threshold), float(slope_obsacle_threshold),
                         float(robot_height), float(robot_radius), float(ground_to_lidar_height),
                         int(xy_eigen_dist), int(z_eigen_dist))
        self.xy_size, self.z_size = int(xy_size), int(z_size)
        self.voxel_count = self.xy_size * self.xy_size * self.z_size
        self.buffer_size = int(buffer_size)
        self.slots: list[ScanSlot | None] = [None] * self.buffer_size
        self.buffer_index = 0
        self.last_buffer_index = 0
        self.ego_position = np.zeros(3)
        self.last: CombinedMap | None = None
        self.combined: CombinedMap | None = None
        self._scratch_hit = np.zeros(self.voxel_count, np.int32)
        self._scratch_total = np.zeros(self.voxel_count, np.int32)

    # -- scan stage (gvom.py:105-220) --------------------------------------------------
    def Process_pointcloud(self, pointcloud, ego_position, transform=None):
        pts = np.ascontiguousarray(pointcloud)
        if pts.dtype not in (np.float32, np.float64):
            pts = pts.astype(np.float64)
        pts = pts.reshape(-1, 3)
        n = pts.shape[0]
        ego = np.ascontiguousarray(np.asarray(ego_position, dtype=np.float64).reshape(3))
        self.ego_position = ego.copy()
        T = None
        tcode = -1
        if transform is not None:
            T = np.ascontiguousarray(transform)
            if T.dtype not in (np.float32, np.float64):
                T = T.astype(np.float64)
            tcode = _dtype_code(T)
        V = self.voxel_count
        cap = max(1, min(V, n))
        index_map = np.empty(V, np.int32)
        hit = np.empty(cap, np.int32)
        total = np.empty(cap, np.int32)
        metrics = np.empty((cap, 10), np.float64)
        min_h = np.empty(cap, np.float32)
        origin = np.zeros(3)
        world = np.empty_like(pts)
        valid = np.empty(max(n, 1), np.uint8)
        cells = self.lib.orc_process(ctypes.byref(self.p), _ptr(pts), n, _dtype_code(pts), _ptr(ego),
                                     _ptr(T), tcode, _ptr(world), _ptr(origin), _ptr(index_map),
                                     _ptr(hit), _ptr(total), _ptr(metrics), _ptr(min_h), cap,
                                     _ptr(self._scratch_hit), _ptr(self._scratch_total), _ptr(valid))
        if cells < 0:
            raise RuntimeError("oracle: cell capacity exceeded")
        slot = ScanSlot(index_map, hit[:cells].copy(), total[:cells].copy(), metrics[:cells].copy(),
                        min_h[:cells].copy(), origin, world, valid[:n].astype(bool))
        self.slots[self.buffer_index] = slot
        self.last_buffer_index = self.buffer_index
        self.buffer_index += 1
        if self.buffer_index >= self.buffer_size:
            self.buffer_index = 0

    process_pointcloud = Process_pointcloud

    # -- combine stage (gvom.py:222-393) -----------------------------------------------
    def combine_maps(self):
        newest = self.slots[self.last_buffer_index]
        if newest is None:
            print("ERROR: No data in buffer")
            return None
        co = newest.origin.copy()
        filled = [s for s in self.slots if s is not None]
        srcs = (_Source * max(1, len(filled)))()
        for k, s in enumerate(filled):
            srcs[k] = _Source(_ptr(s.index_map).value, _ptr(s.hit).value, _ptr(s.total).value,
                              _ptr(s.min_h).value, _ptr(s.metrics).value, _ptr(s.origin).value)
        last = None
        cap = sum(len(s.hit) for s in filled)
        if self.last is not None:
            L = self.last
            last = _Source(_ptr(L.index_map).value, _ptr(L.hit).value, _ptr(L.total).value,
                           _ptr(L.min_h).value, _ptr(L.metrics).value, _ptr(L.origin).value)
            cap += len(L.hit)
        cap = max(cap, 1)
        V, XY = self.voxel_count, self.xy_size
        comb = np.empty(V, np.int32)
        chit = np.empty(cap, np.int32)
        ctot = np.empty(cap, np.int32)
        cmh = np.empty(cap, np.float32)
        cmet = np.empty((cap, 10), np.float32)
        ev = np.empty((cap, 3), np.float32)
        maps_f = [np.empty((XY, XY), np.float64) for _ in range(6)]
        maps_i = [np.empty((XY, XY), np.int32) for _ in range(3)]
        cells = self.lib.orc_combine(ctypes.byref(self.p), len(filled), srcs,
                                     None if last is None else ctypes.byref(last), _ptr(co),
                                     _ptr(self.ego_position), _ptr(comb), _ptr(chit), _ptr(ctot),
                                     _ptr(cmh), _ptr(cmet), _ptr(ev), cap,
                                     *[_ptr(m) for m in maps_f], *[_ptr(m) for m in maps_i])
        if cells < 0:
            raise RuntimeError("oracle: combined capacity exceeded")
        height, inferred, sx, sy, rough, delta = maps_f
        pos, neg, vis = maps_i
        cm = CombinedMap(comb, chit[:cells].copy(), ctot[:cells].copy(), cmh[:cells].copy(),
                         cmet[:cells].copy(), ev[:cells].copy(), co, height, inferred, sx, sy,
                         rough, delta, pos, neg, vis)
        self.combined = cm
        self.last = cm
        origin_m = co.copy()
        origin_m[0] *= self.p.xy_resolution
        origin_m[1] *= self.p.xy_resolution
        origin_m[2] *= self.p.z_resolution
        return (origin_m, pos.copy(), neg.copy(), rough.copy(), vis.copy())

    # -- debug exporters (gvom.py:395-442, kernels 455-503) ----------------------------
    def make_debug_voxel_map(self):
        if self.combined is None:
            print("No data")
            return None
        return debug_voxel_map(self.combined, self.xy_size, self.z_size,
                               self.p.xy_resolution, self.p.z_resolution)

    def make_debug_height_map(self):
        if self.combined is None:
            print("No data")
            return None
        return debug_height_map(self.combined, self.xy_size, self.p.xy_resolution, self.p.z_resolution)

    def make_debug_inferred_height_map(self):
        if self.combined is None:
            print("No data")
            return None
        return debug_inferred_height_map(self.combined, self.xy_size, self.p.xy_resolution,
                                         self.p.z_resolution)


def debug_voxel_map(cm: CombinedMap, xy, z, xyr, zr):
    """__make_voxel_pointcloud (gvom.py:482-503): one f32 row of 8 per combined cell id."""
    C = len(cm.hit)
    out = np.zeros((C, 8), np.float32)
    idx = cm.index_map
    v = np.nonzero(idx >= 0)[0]
    ids = idx[v]
    x = v % xy
    y = (v // xy) % xy
    zz = v // (xy * xy)
    out[ids, 0] = (x + cm.origin[0]) * xyr
    out[ids, 1] = (y + cm.origin[1]) * xyr
    out[ids, 2] = (zz + cm.origin[2]) * zr
    if _PLANTED == 6:
        out[ids, 3] = cm.total[ids].astype(np.float64) / cm.hit[ids].astype(np.float64)
    else:
        out[ids, 3] = cm.hit[ids].astype(np.float64) / cm.total[ids].astype(np.float64)
    out[ids, 4] = cm.hit[ids]
    e = cm.eigen[ids]
    out[ids, 5] = e[:, 0] - (e[:, 2] if _PLANTED == 7 else e[:, 1])
    out[ids, 6] = e[:, 1] - e[:, 2]
    out[ids, 7] = e[:, 2]
    return out


def debug_height_map(cm: CombinedMap, xy, xyr, zr):
    """__make_height_map_pointcloud (gvom.py:456-468): row index x + y*xy."""
    out = np.zeros((xy * xy, 7), np.float32)
    x, y = np.meshgrid(np.arange(xy), np.arange(xy), indexing="ij")
    index = (x * xy + y if _PLANTED == 8 else x + y * xy).ravel()
    out[index, 0] = ((x + cm.origin[0]) * xyr).ravel()
    out[index, 1] = ((y + cm.origin[1]) * xyr).ravel()
    out[index, 2] = (cm.height - zr).ravel()
    out[index, 3] = cm.rough.ravel()
    out[index, 4] = cm.slope_x.ravel()
    out[index, 5] = cm.slope_y.ravel()
    out[index, 6] = np.sqrt(cm.slope_x * cm.slope_x + cm.slope_y * cm.slope_y).ravel()
    return out


def debug_inferred_height_map(cm: CombinedMap, xy, xyr, zr):
    """__make_infered_height_map_pointcloud (gvom.py:471-479), called with the guessed
    height delta as its height input (gvom.py:439-440)."""
    out = np.zeros((xy * xy, 3), np.float32)
    x, y = np.meshgrid(np.arange(xy), np.arange(xy), indexing="ij")
    index = (x * xy + y if _PLANTED == 8 else x + y * xy).ravel()
    out[index, 0] = ((x + cm.origin[0]) * xyr).ravel()
    out[index, 1] = ((y + cm.origin[1]) * xyr).ravel()
    out[index, 2] = (cm.delta - zr).ravel()
    return out


def eigenvalues(metrics_f32):
    """The oracle's __calculate_eigenvalues on rows of a f32 [C, 10] metrics array."""
    lib = load_library()
    m = np.ascontiguousarray(metrics_f32, dtype=np.float32).reshape(-1, 10)
    ev = np.empty((m.shape[0], 3), np.float32)
    lib.orc_eigenvalues(_ptr(m), _ptr(ev), m.shape[0])
    return ev


def slope_guess(height, inferred, xy_resolution):
    """The oracle's __calculate_slope + __guess_height on [x, y] maps -> (sx, sy, rough, delta)."""
    lib = load_library()
    h = np.ascontiguousarray(height, dtype=np.float64)
    f = np.ascontiguousarray(inferred, dtype=np.float64)
    XY = h.shape[0]
    p = _Params(float(xy_resolution), 1.0, XY, 4, 1, 0, 1.0, 0.5, 0.5, 0.3, 2.0, 4.0, 1.0, 1, 1)
    outs = [np.empty((XY, XY), np.float64) for _ in range(4)]
    lib.orc_slope_guess(ctypes.byref(p), _ptr(h), _ptr(f), *[_ptr(o) for o in outs])
    return tuple(outs)


# -- canonicalisation ------------------------------------------------------------------
def canonical_scan(index_map, hit, total, min_h=None, metrics=None):
    """Per-voxel dense fields of one index map + compact arrays, independent of cell-id order.

    Returns a dict of arrays indexed by reference linear voxel id:
      occ (bool), passes (int64: t for free voxels, 0 otherwise), hit, total (int64),
      min_h (f32, NaN where unoccupied), metrics (f64[V,10], NaN where unoccupied).
    """
    idx = np.asarray(index_map)
    occ = idx >= 0
    V = idx.shape[0]
    passes = np.where(idx < -1, -(idx.astype(np.int64)) - 1, 0)
    hd = np.zeros(V, np.int64)
    td = np.zeros(V, np.int64)
    ids = idx[occ]
    hd[occ] = np.asarray(hit)[ids]
    td[occ] = np.asarray(total)[ids]
    out = {"occ": occ, "passes": passes, "hit": hd, "total": td, "cells": int(occ.sum())}
    if min_h is not None:
        mh = np.full(V, np.nan, np.float32)
        mh[occ] = np.asarray(min_h)[ids]
        out["min_h"] = mh
    if metrics is not None:
        m = np.asarray(metrics)
        md = np.full((V, m.shape[1] if m.ndim == 2 else 10), np.nan, np.float64)
        md[occ] = m[ids]
        out["metrics"] = md
    return out


def origin_of(ego, xy_resolution, z_resolution, xy_size, z_size):
    """gvom.py:138-141 in Python floats."""
    import math
    return np.array([math.floor((ego[0] / xy_resolution) - xy_size / 2),
                     math.floor((ego[1] / xy_resolution) - xy_size / 2),
                     math.floor((ego[2] / z_resolution) - z_size / 2)], np.float64)
