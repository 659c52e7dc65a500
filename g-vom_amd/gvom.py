"""Drop-in `Gvom` for G-VOM's scan->map path on MI355X.

Same constructor (14 positional arguments, gvom.py:21-22), the same methods
`Process_pointcloud` (gvom.py:105; also exported as `process_pointcloud`, the name the
reference README uses), `combine_maps` (gvom.py:222) and the three debug exporters
(gvom.py:395-442), with the same return values.  All compute runs in
libgvom_hip.so (hand-written gfx950 kernels) through its C-ABI; this class only
converts arguments and allocates the host arrays it returns.

Differences from the reference, all in cases the reference leaves undefined or
racy:
  - `ego_position` is copied, not aliased (gvom.py:111 stores the caller's object);
  - points with a non-finite coordinate are dropped (the reference indexes with NaN);
  - an empty scan, or a scan or map without occupied cells, gives the empty-state
    result instead of a zero-block launch error;
  - concurrent calls are serialised per instance (the reference advances
    buffer_index without a lock, gvom.py:211-216).
"""
from __future__ import annotations

import ctypes

import numpy as np

import gvom_native as nat


def _as_points(pointcloud):
    pts = np.asarray(pointcloud)
    if pts.dtype not in (np.float32, np.float64):
        pts = pts.astype(np.float64)
    pts = np.ascontiguousarray(pts).reshape(-1, 3)
    return pts, (nat.GVOM_F32 if pts.dtype == np.float32 else nat.GVOM_F64)


def _as_transform(transform):
    if transform is None:
        return None, nat.GVOM_NO_TRANSFORM
    T = np.asarray(transform)
    if T.dtype not in (np.float32, np.float64):
        T = T.astype(np.float64)
    T = np.ascontiguousarray(T)
    if T.shape != (4, 4):
        raise ValueError(f"transform must be 4x4, got {T.shape}")
    return T, (nat.GVOM_F32 if T.dtype == np.float32 else nat.GVOM_F64)


def cloud_layout(cloud):
    """gvom_cloud_layout of a PointCloud2-like message: the x, y, z PointFields."""
    fields = {f.name: f for f in cloud.fields}
    for name in ("x", "y", "z"):
        if name not in fields:
            raise ValueError(f"PointCloud2 has no '{name}' field")
    lay = nat.CloudLayout()
    lay.height = int(cloud.height)
    lay.width = int(cloud.width)
    lay.point_step = int(cloud.point_step)
    lay.row_step = int(cloud.row_step)
    for k, name in enumerate(("x", "y", "z")):
        lay.offset[k] = int(fields[name].offset)
        lay.datatype[k] = int(fields[name].datatype)
    lay.is_bigendian = 1 if cloud.is_bigendian else 0
    return lay


class Gvom:
    """Scan -> voxel map -> 2D cost maps (reference class: scripts/gvom.py:8)."""

    def __init__(self, xy_resolution, z_resolution, xy_size, z_size, buffer_size, min_distance,
                 positive_obstacle_threshold, negative_obstacle_threshold, slope_obsacle_threshold,
                 robot_height, robot_radius, ground_to_lidar_height, xy_eigen_dist, z_eigen_dist):
        self._lib = nat.load()
        self.xy_resolution = xy_resolution
        self.z_resolution = z_resolution
        self.xy_size = int(xy_size)
        self.z_size = int(z_size)
        self.buffer_size = int(buffer_size)
        self.min_distance = min_distance
        self.positive_obstacle_threshold = positive_obstacle_threshold
        self.negative_obstacle_threshold = negative_obstacle_threshold
        self.slope_obsacle_threshold = slope_obsacle_threshold
        self.robot_height = robot_height
        self.robot_radius = robot_radius
        self.ground_to_lidar_height = ground_to_lidar_height
        self.xy_eigen_dist = int(xy_eigen_dist)
        self.z_eigen_dist = int(z_eigen_dist)
        self.metrics_count = 10
        self.voxel_count = self.xy_size * self.xy_size * self.z_size
        self.ego_position = [0, 0, 0]
        p = nat.Params(float(xy_resolution), float(z_resolution), self.xy_size, self.z_size,
                       self.buffer_size, 0, float(min_distance), float(positive_obstacle_threshold),
                       float(negative_obstacle_threshold), float(slope_obsacle_threshold),
                       float(robot_height), float(robot_radius), float(ground_to_lidar_height),
                       self.xy_eigen_dist, self.z_eigen_dist)
        h = ctypes.c_void_p()
        nat.check(self._lib.gvom_create(ctypes.byref(p), ctypes.byref(h)), None)
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.gvom_destroy(h)
            self._h = None

    # -- gvom.py:105-220 -------------------------------------------------------------
    def Process_pointcloud(self, pointcloud, ego_position, transform=None):
        """Imports a pointcloud, processes it into a voxel map and adds it to the buffer."""
        pts, dt = _as_points(pointcloud)
        ego = np.ascontiguousarray(np.asarray(ego_position, dtype=np.float64).reshape(3))
        T, tdt = _as_transform(transform)
        self.ego_position = tuple(float(v) for v in ego)
        nat.check(self._lib.gvom_process_pointcloud(self._h, nat.ptr(pts), pts.shape[0], dt, nat.ptr(ego),
                                                    nat.ptr(T), tdt), self._h)

    process_pointcloud = Process_pointcloud

    def process_pointcloud_device(self, dev_points, n, ego_position, transform=None, dtype=np.float32):
        """Process_pointcloud on points already in device memory (dev_points: a device
        pointer as int / c_void_p / DeviceArray).  Does not synchronise."""
        ptr = dev_points.ptr if isinstance(dev_points, nat.DeviceArray) else ctypes.c_void_p(int(dev_points))
        ego = np.ascontiguousarray(np.asarray(ego_position, dtype=np.float64).reshape(3))
        T, tdt = _as_transform(transform)
        dt = nat.GVOM_F32 if np.dtype(dtype) == np.float32 else nat.GVOM_F64
        self.ego_position = tuple(float(v) for v in ego)
        nat.check(self._lib.gvom_process_pointcloud_device(self._h, ptr, int(n), dt, nat.ptr(ego), nat.ptr(T), tdt),
                  self._h)

    def process_pointcloud2(self, cloud, ego_position, transform=None, device_data=None):
        """The ROS node's input path (gvom_ros.py:108-109) on a sensor_msgs/PointCloud2
        (or any object with height, width, point_step, row_step, is_bigendian, fields
        (name, offset, datatype) and data): x, y, z are read from the message bytes on the
        device, as float64, and non-finite points are dropped, the same slot that
        Process_pointcloud(ros_numpy.point_cloud2.pointcloud2_to_xyz_array(cloud), ...)
        makes.  device_data: the message bytes already in device memory (then
        cloud.data is not read and the call does not synchronise)."""
        lay = cloud_layout(cloud)
        ego = np.ascontiguousarray(np.asarray(ego_position, dtype=np.float64).reshape(3))
        T, tdt = _as_transform(transform)
        self.ego_position = tuple(float(v) for v in ego)
        if device_data is not None:
            ptr = device_data.ptr if isinstance(device_data, nat.DeviceArray) else ctypes.c_void_p(int(device_data))
            on_dev = 1
            keep = None
        else:
            keep = np.frombuffer(memoryview(cloud.data).cast("B"), dtype=np.uint8)
            need = (lay.height - 1) * lay.row_step + lay.width * lay.point_step if lay.height * lay.width else 0
            if keep.size < need:
                raise ValueError(f"PointCloud2 data holds {keep.size} bytes, layout needs {need}")
            ptr = keep.ctypes.data_as(ctypes.c_void_p)
            on_dev = 0
        nat.check(self._lib.gvom_process_pointcloud2(self._h, ptr, ctypes.byref(lay), nat.ptr(ego), nat.ptr(T), tdt,
                                                     on_dev), self._h)
        del keep

    def reserve(self, max_points, max_batch_points=0):
        nat.check(self._lib.gvom_reserve(self._h, int(max_points), int(max_batch_points)), self._h)

    def process_batch(self, points, offsets, egos, transforms=None):
        """Several scans in one launch sequence; same result as len(offsets)-1
        consecutive Process_pointcloud calls.  points: flat (sum N, 3) array."""
        pts, dt = _as_points(points)
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        ego = np.ascontiguousarray(np.asarray(egos, dtype=np.float64).reshape(-1, 3))
        S = off.shape[0] - 1
        if ego.shape[0] != S:
            raise ValueError("one ego position per scan is required")
        T, tdt = None, nat.GVOM_NO_TRANSFORM
        if transforms is not None:
            T = np.asarray(transforms)
            if T.dtype not in (np.float32, np.float64):
                T = T.astype(np.float64)
            T = np.ascontiguousarray(T).reshape(S, 4, 4)
            tdt = nat.GVOM_F32 if T.dtype == np.float32 else nat.GVOM_F64
        self.ego_position = tuple(float(v) for v in ego[-1])
        nat.check(self._lib.gvom_process_batch(self._h, nat.ptr(pts), nat.ptr(off), S, dt, nat.ptr(ego),
                                               nat.ptr(T), tdt, 0), self._h)

    def process_batch_device(self, dev_points, offsets, egos, transforms=None, dtype=np.float32):
        """process_batch on a flat point array already in device memory (pointer as int,
        c_void_p or DeviceArray).  Does not synchronise; capturable in a graph."""
        ptr = dev_points.ptr if isinstance(dev_points, nat.DeviceArray) else ctypes.c_void_p(int(dev_points))
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        ego = np.ascontiguousarray(np.asarray(egos, dtype=np.float64).reshape(-1, 3))
        S = off.shape[0] - 1
        T, tdt = None, nat.GVOM_NO_TRANSFORM
        if transforms is not None:
            T = np.asarray(transforms)
            if T.dtype not in (np.float32, np.float64):
                T = T.astype(np.float64)
            T = np.ascontiguousarray(T).reshape(S, 4, 4)
            tdt = nat.GVOM_F32 if T.dtype == np.float32 else nat.GVOM_F64
        dt = nat.GVOM_F32 if np.dtype(dtype) == np.float32 else nat.GVOM_F64
        self.ego_position = tuple(float(v) for v in ego[-1])
        nat.check(self._lib.gvom_process_batch(self._h, ptr, nat.ptr(off), S, dt, nat.ptr(ego), nat.ptr(T), tdt, 1),
                  self._h)

    def autotune(self, dev_points, n, ego_position, transform=None, dtype=np.float32):
        """Pick the ray-cast variant by timing it on one device-resident scan.  Returns
        (block, rays_per_thread, flags, best_ms); flags bit 0 LDS box, bit 1 sector order,
        bit 2 tiled scratch, bit 3 spread (kept as set, not searched)."""
        ptr = dev_points.ptr if isinstance(dev_points, nat.DeviceArray) else ctypes.c_void_p(int(dev_points))
        ego = np.ascontiguousarray(np.asarray(ego_position, dtype=np.float64).reshape(3))
        T, tdt = _as_transform(transform)
        dt = nat.GVOM_F32 if np.dtype(dtype) == np.float32 else nat.GVOM_F64
        chosen = (ctypes.c_int32 * 3)()
        best = ctypes.c_float()
        nat.check(self._lib.gvom_autotune(self._h, ptr, int(n), dt, nat.ptr(ego), nat.ptr(T), tdt, chosen,
                                          ctypes.byref(best)), self._h)
        return chosen[0], chosen[1], chosen[2], best.value

    def probe_atomics(self, pattern=0, n_atomics=1 << 24):
        """Time non-returning int32 atomic adds (pattern: 0 scattered, 1 64 hot words,
        2 scattered within the issuing XCD's region, 3 scattered at workgroup scope, 4 a
        wavefront's lanes on 64 consecutive words, 5 a wavefront on 8 lines of 8 lanes).
        Returns (ms, atomics issued)."""
        out = (ctypes.c_float * 2)()
        nat.check(self._lib.gvom_probe_atomics(self._h, int(pattern), int(n_atomics), out), self._h)
        return out[0], int(out[1])

    def set_raycast_config(self, block=128, rays_per_thread=1, use_box=False, sector_order=False, tiled=True,
                           spread=True):
        """Ray-cast launch shape and variant (tiled scratch excludes the LDS box; spread takes
        each block's wavefronts from across the scan); a new handle starts at the defaults."""
        flags = (1 if use_box else 0) | (2 if sector_order else 0) | (4 if tiled else 0) | (8 if spread else 0)
        nat.check(self._lib.gvom_set_raycast_config(self._h, int(block), int(rays_per_thread), flags), self._h)

    def set_ray_rings(self, points_per_ring, patch_rings=32):
        """Ray order hint for organised scans (rings of points_per_ring consecutive points):
        each wavefront of the ray cast traces 64 / patch_rings azimuths x patch_rings rings.
        Same results for any order; 0 restores the linear order."""
        nat.check(self._lib.gvom_set_ray_rings(self._h, int(points_per_ring), int(patch_rings)), self._h)

    def set_overlap(self, enabled=True):
        """Let a device-pointer process call run while the preceding combine_maps_device
        still runs (off by default).  Its points must be complete when the call is made."""
        nat.check(self._lib.gvom_set_overlap(self._h, 1 if enabled else 0), self._h)

    # -- hipGraph capture ------------------------------------------------------------
    def graph_begin(self, max_scans=64):
        """Start recording device-pointer process calls and combine_maps_device calls."""
        nat.check(self._lib.gvom_graph_begin(self._h, int(max_scans)), self._h)

    def graph_end(self):
        g = ctypes.c_void_p()
        nat.check(self._lib.gvom_graph_end(self._h, ctypes.byref(g)), self._h)
        return GvomGraph(self, g)

    # -- gvom.py:222-393 -------------------------------------------------------------
    def combine_maps(self):
        """Combines all maps in the buffer and processes them into 2D maps."""
        X = self.xy_size
        origin = np.zeros(3, np.float64)
        pos = np.empty((X, X), np.int32)
        neg = np.empty((X, X), np.int32)
        rough = np.empty((X, X), np.float64)
        vis = np.empty((X, X), np.int32)
        rc = nat.check(self._lib.gvom_combine_maps(self._h, nat.ptr(origin), nat.ptr(pos), nat.ptr(neg),
                                                   nat.ptr(rough), nat.ptr(vis)), self._h)
        if rc == nat.GVOM_NO_DATA:
            print("ERROR: No data in buffer")
            return None
        return (origin, pos, neg, rough, vis)

    def combine_maps_device(self):
        """combine_maps without host copies or synchronisation; returns the origin in
        metres, or None when the buffer is empty."""
        origin = np.zeros(3, np.float64)
        rc = nat.check(self._lib.gvom_combine_maps_device(self._h, nat.ptr(origin)), self._h)
        return None if rc == nat.GVOM_NO_DATA else origin

    # -- gvom_ros.py:137-164 ---------------------------------------------------------
    def occupancy_grids(self, density_threshold=50, min_roughness=-10, max_roughness=0, device_out=None):
        """OccupancyGrid payloads of the latest combined maps, as the ROS node publishes
        them (gvom_ros.py:137-164): a dict topic -> int8 array of xy_size^2 bytes in the
        message's order (x fastest), plus "origin" (metres; the node publishes x, y and
        z = 0).  Topics: hard_obstacle_map, soft_obstacle_map, ground_certainty_map (the
        node also publishes it as all_ground_certainty_map), negative_obstacle_map,
        roughness_map.  With device_out (a device pointer to 5 * xy_size^2 bytes) the
        payloads are written there without synchronising and only the origin is
        returned.  None before the first combine."""
        X = self.xy_size * self.xy_size
        origin = np.zeros(3, np.float64)
        if device_out is not None:
            ptr = device_out.ptr if isinstance(device_out, nat.DeviceArray) else ctypes.c_void_p(int(device_out))
            rc = nat.check(self._lib.gvom_occupancy_grids(self._h, float(density_threshold), float(min_roughness),
                                                          float(max_roughness), ptr, 1, nat.ptr(origin)), self._h)
            return None if rc == nat.GVOM_NO_DATA else origin
        out = np.empty(5 * X, np.int8)
        rc = nat.check(self._lib.gvom_occupancy_grids(self._h, float(density_threshold), float(min_roughness),
                                                      float(max_roughness), nat.ptr(out), 0, nat.ptr(origin)),
                       self._h)
        if rc == nat.GVOM_NO_DATA:
            return None
        grids = {name: out[k * X:(k + 1) * X] for k, name in enumerate(nat.GRID_NAMES)}
        grids["origin"] = origin
        return grids

    # -- gvom.py:395-442 -------------------------------------------------------------
    def make_debug_voxel_map(self):
        rows = ctypes.c_int64(0)
        rc = nat.check(self._lib.gvom_debug_voxel_map(self._h, None, 0, ctypes.byref(rows)), self._h)
        if rc == nat.GVOM_NO_DATA:
            print("No data")
            return None
        out = np.zeros((rows.value, 8), np.float32)
        nat.check(self._lib.gvom_debug_voxel_map(self._h, nat.ptr(out), rows.value, ctypes.byref(rows)), self._h)
        return out

    def make_debug_height_map(self):
        out = np.zeros((self.xy_size * self.xy_size, 7), np.float32)
        rc = nat.check(self._lib.gvom_debug_height_map(self._h, nat.ptr(out)), self._h)
        if rc == nat.GVOM_NO_DATA:
            print("No data")
            return None
        return out

    def make_debug_inferred_height_map(self):
        out = np.zeros((self.xy_size * self.xy_size, 3), np.float32)
        rc = nat.check(self._lib.gvom_debug_inferred_height_map(self._h, nat.ptr(out)), self._h)
        if rc == nat.GVOM_NO_DATA:
            print("No data")
            return None
        return out

    # -- state export (no reference method; the reference keeps these as members) ---
    def ring_state(self):
        vals = [ctypes.c_int32() for _ in range(4)]
        nat.check(self._lib.gvom_ring_state(self._h, *[ctypes.byref(v) for v in vals]), self._h)
        return dict(buffer_index=vals[0].value, last_buffer_index=vals[1].value,
                    filled_mask=vals[2].value, have_combined=bool(vals[3].value))

    def slot_state(self, slot):
        """Slot `slot` in the reference's format: (index_map int32[V], hit, total, min_h,
        metrics f64[C,10], origin f64[3]) or None when the slot is empty."""
        cells = ctypes.c_int64(0)
        origin = np.zeros(3)
        rc = nat.check(self._lib.gvom_get_slot(self._h, slot, None, 0, None, None, None, None, nat.ptr(origin),
                                               ctypes.byref(cells)), self._h)
        if rc == nat.GVOM_NO_DATA:
            return None
        C = cells.value
        idx = np.empty(self.voxel_count, np.int32)
        hit = np.empty(C, np.int32)
        tot = np.empty(C, np.int32)
        mh = np.empty(C, np.float32)
        met = np.empty((C, 10), np.float64)
        nat.check(self._lib.gvom_get_slot(self._h, slot, nat.ptr(idx), C, nat.ptr(hit), nat.ptr(tot), nat.ptr(mh),
                                          nat.ptr(met), nat.ptr(origin), ctypes.byref(cells)), self._h)
        return dict(index_map=idx, hit=hit, total=tot, min_h=mh, metrics=met, origin=origin)

    def combined_state(self):
        cells = ctypes.c_int64(0)
        origin = np.zeros(3)
        rc = nat.check(self._lib.gvom_get_combined(self._h, None, 0, None, None, None, None, None, nat.ptr(origin),
                                                   ctypes.byref(cells)), self._h)
        if rc == nat.GVOM_NO_DATA:
            return None
        C = cells.value
        idx = np.empty(self.voxel_count, np.int32)
        hit = np.empty(C, np.int32)
        tot = np.empty(C, np.int32)
        mh = np.empty(C, np.float32)
        met = np.empty((C, 10), np.float32)
        ev = np.empty((C, 3), np.float32)
        nat.check(self._lib.gvom_get_combined(self._h, nat.ptr(idx), C, nat.ptr(hit), nat.ptr(tot), nat.ptr(mh),
                                              nat.ptr(met), nat.ptr(ev), nat.ptr(origin), ctypes.byref(cells)),
                  self._h)
        return dict(index_map=idx, hit=hit, total=tot, min_h=mh, metrics=met, eigen=ev, origin=origin)

    def maps_state(self):
        X = self.xy_size
        f = {k: np.empty((X, X), np.float64) for k in ("height", "inferred", "slope_x", "slope_y", "rough", "delta")}
        i = {k: np.empty((X, X), np.int32) for k in ("pos", "neg", "vis")}
        rc = nat.check(self._lib.gvom_get_maps(self._h, *[nat.ptr(f[k]) for k in
                                                           ("height", "inferred", "slope_x", "slope_y", "rough",
                                                            "delta")],
                                               *[nat.ptr(i[k]) for k in ("pos", "neg", "vis")]), self._h)
        if rc == nat.GVOM_NO_DATA:
            return None
        return {**f, **i}

    def set_timing(self, enabled=True):
        nat.check(self._lib.gvom_set_timing(self._h, 1 if enabled else 0), self._h)

    def stage_times(self):
        ms = (ctypes.c_float * 9)()
        nat.check(self._lib.gvom_stage_times(self._h, ms, 9), self._h)
        return list(ms)

    # -- gvom.py:807-885: image kernels the reference defines but never calls -----------
    def expand_binary(self, img, r):
        """__expand_binary (gvom.py:807-832): disk dilation of a 0/1 int32 map, radius r."""
        img = np.ascontiguousarray(np.asarray(img, dtype=np.int32))
        X = self.xy_size
        if img.shape != (X, X):
            raise ValueError(f"image must be {X} x {X}")
        din, dout = nat.DeviceArray(img), nat.DeviceArray(np.zeros_like(img))
        nat.check(self._lib.gvom_expand_binary(self._h, din.ptr, dout.ptr, X, float(r)), self._h)
        self.sync()
        return dout.to_host()

    def lowpass_binary(self, img, filter_size, filter_fraction):
        """__lowpass_binary (gvom.py:834-862) on a 0/1 int32 map."""
        img = np.ascontiguousarray(np.asarray(img, dtype=np.int32))
        X = self.xy_size
        if img.shape != (X, X):
            raise ValueError(f"image must be {X} x {X}")
        din, dout = nat.DeviceArray(img), nat.DeviceArray(np.zeros_like(img))
        nat.check(self._lib.gvom_lowpass_binary(self._h, din.ptr, dout.ptr, X, int(filter_size),
                                                float(filter_fraction)), self._h)
        self.sync()
        return dout.to_host()

    def convolve(self, img, kernel):
        """__convolve (gvom.py:864-885) with its arithmetic as written (every term reads
        img[x + r, y + r]); kernel: odd k x k float64."""
        img = np.ascontiguousarray(np.asarray(img, dtype=np.float64))
        kernel = np.ascontiguousarray(np.asarray(kernel, dtype=np.float64))
        X = self.xy_size
        if img.shape != (X, X) or kernel.ndim != 2 or kernel.shape[0] != kernel.shape[1]:
            raise ValueError(f"image must be {X} x {X} and the kernel square")
        din, dk, dout = nat.DeviceArray(img), nat.DeviceArray(kernel), nat.DeviceArray(np.zeros_like(img))
        nat.check(self._lib.gvom_convolve(self._h, din.ptr, dk.ptr, dout.ptr, X, kernel.shape[0]), self._h)
        self.sync()
        return dout.to_host()

    def sync(self):
        nat.check(self._lib.gvom_sync(self._h), self._h)

    @property
    def stream(self):
        """The handle's HIP stream (every call's work is ordered on it), as c_void_p."""
        return ctypes.c_void_p(self._lib.gvom_stream(self._h))


class GvomGraph:
    """A captured launch sequence; launch() replays it with new ego positions and
    transforms for the captured scans (capture order)."""

    def __init__(self, owner: Gvom, handle):
        self.owner = owner
        self._g = handle
        self.scans = owner._lib.gvom_graph_scans(handle)

    def launch(self, egos, transforms=None):
        ego = np.ascontiguousarray(np.asarray(egos, dtype=np.float64).reshape(-1, 3))
        if ego.shape[0] != self.scans:
            raise ValueError(f"graph holds {self.scans} scans, got {ego.shape[0]} ego positions")
        T, tdt = None, nat.GVOM_NO_TRANSFORM
        if transforms is not None:
            T = np.asarray(transforms)
            if T.dtype not in (np.float32, np.float64):
                T = T.astype(np.float64)
            T = np.ascontiguousarray(T).reshape(self.scans, 4, 4)
            tdt = nat.GVOM_F32 if T.dtype == np.float32 else nat.GVOM_F64
        nat.check(self.owner._lib.gvom_graph_launch(self.owner._h, self._g, nat.ptr(ego), nat.ptr(T), tdt),
                  self.owner._h)

    def __del__(self):
        g = getattr(self, "_g", None)
        if g is not None and g.value:
            self.owner._lib.gvom_graph_destroy(g)
            self._g = None
