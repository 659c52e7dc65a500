"""Seeded synthetic spinning-lidar scans over procedural off-road terrain.

The workloads are the scan->map configurations this package is tested and
benchmarked on (there is no dataset on the box): a heightfield of 8 random 2D
sinusoids plus 3 Gaussian bumps, axis-aligned boxes as positive obstacles, pits as
negative obstacles and, optionally, vegetation volumes.  A spinning lidar with
evenly spaced rings and uniform azimuths is ray-marched against it in float64,
ranges get N(0, 0.02 m) noise and dropped returns are removed, so scans are ragged.
Points are emitted ring-major (all azimuths of ring 0, then ring 1, ...), the order
an organised PointCloud2 from a spinning lidar has.

Rays that hit no modelled surface within the range clip return from a far surface
at a uniformly drawn range in [r_max/2, r_max], so a scan keeps its nominal size
before dropout.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Map parameters shared by all workloads (ROS defaults, scripts/gvom_ros.py:23-41).
DEFAULTS = dict(min_distance=1.0, positive_obstacle_threshold=0.5, negative_obstacle_threshold=0.5,
                slope_obsacle_threshold=0.3, robot_height=2.0, robot_radius=4.0,
                ground_to_lidar_height=1.0, xy_eigen_dist=1, z_eigen_dist=1)


@dataclass
class World:
    amp: np.ndarray
    lam: np.ndarray
    theta: np.ndarray
    phi: np.ndarray
    bumps: np.ndarray        # (3, 4): cx, cy, amplitude, sigma
    boxes: np.ndarray        # (nb, 5): x0, x1, y0, y1, height above terrain
    pits: np.ndarray         # (np, 5): x0, x1, y0, y1, depth
    veg: np.ndarray          # (nv, 3): cx, cy, radius (2 m tall volumes over the terrain)
    h0: float = 0.0          # offset so that the terrain height at (0, 0) is 0

    def terrain(self, x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        h = np.zeros(np.broadcast(x, y).shape)
        for a, l, t, p in zip(self.amp, self.lam, self.theta, self.phi):
            h += a * np.sin(2 * np.pi * (x * np.cos(t) + y * np.sin(t)) / l + p)
        for cx, cy, a, s in self.bumps:
            h += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
        for x0, x1, y0, y1, d in self.pits:
            h -= d * ((x >= x0) & (x <= x1) & (y >= y0) & (y <= y1))
        return h - self.h0

    def surface(self, x, y):
        """Terrain plus boxes (boxes stand on the terrain)."""
        h = self.terrain(x, y)
        for x0, x1, y0, y1, bh in self.boxes:
            inside = (x >= x0) & (x <= x1) & (y >= y0) & (y <= y1)
            h = np.where(inside, h + bh, h)
        return h


def make_world(seed: int, n_boxes: int, n_pits: int, vegetation: bool, extent: float = 60.0) -> World:
    rng = np.random.default_rng(np.random.SeedSequence(seed).spawn(1)[0])
    amp = rng.uniform(0.05, 0.5, 8)
    lam = rng.uniform(2.0, 20.0, 8)
    theta = rng.uniform(0, 2 * np.pi, 8)
    phi = rng.uniform(0, 2 * np.pi, 8)
    bumps = np.stack([rng.uniform(-extent, extent, 3), rng.uniform(-extent, extent, 3),
                      rng.uniform(0.3, 1.5, 3), rng.uniform(2.0, 6.0, 3)], axis=1)

    def rects(n, lo, hi, size_lo, size_hi, val_lo, val_hi):
        out = []
        for _ in range(n):
            while True:
                cx, cy = rng.uniform(lo, hi, 2)
                if math.hypot(cx, cy) > 4.0:   # keep the start pose clear
                    break
            sx, sy = rng.uniform(size_lo, size_hi, 2)
            out.append([cx - sx / 2, cx + sx / 2, cy - sy / 2, cy + sy / 2, rng.uniform(val_lo, val_hi)])
        return np.array(out, np.float64).reshape(-1, 5)

    boxes = rects(n_boxes, -0.6 * extent, 0.6 * extent, 0.6, 3.0, 0.5, 2.0)
    pits = rects(n_pits, -0.6 * extent, 0.6 * extent, 1.5, 4.0, 0.5, 1.5)
    veg = np.zeros((0, 3))
    if vegetation:
        # discs covering ~5 % of a (2*extent)^2 area
        area = 0.05 * (2 * extent) ** 2
        radii = []
        while sum(np.pi * r * r for r in radii) < area:
            radii.append(rng.uniform(2.0, 5.0))
        k = len(radii)
        veg = np.stack([rng.uniform(-extent, extent, k), rng.uniform(-extent, extent, k),
                        np.array(radii)], axis=1)
    w = World(amp, lam, theta, phi, bumps, boxes, pits, veg)
    w.h0 = float(w.terrain(0.0, 0.0))
    return w


def _raster(world: World, x0, x1, y0, y1, res=0.1):
    nx = int(math.ceil((x1 - x0) / res)) + 2
    ny = int(math.ceil((y1 - y0) / res)) + 2
    xs = x0 + res * np.arange(nx)
    ys = y0 + res * np.arange(ny)
    H = np.empty((nx, ny), np.float32)
    step = max(1, 2_000_000 // ny)
    for i in range(0, nx, step):
        X, Y = np.meshgrid(xs[i:i + step], ys, indexing="ij")
        H[i:i + step] = world.surface(X, Y)
    return H, x0, y0, res


def _sample(R, x, y):
    H, x0, y0, res = R
    fx = (x - x0) / res
    fy = (y - y0) / res
    ix = np.clip(np.floor(fx).astype(np.int64), 0, H.shape[0] - 2)
    iy = np.clip(np.floor(fy).astype(np.int64), 0, H.shape[1] - 2)
    tx = np.clip(fx - ix, 0.0, 1.0)
    ty = np.clip(fy - iy, 0.0, 1.0)
    h00 = H[ix, iy]
    h10 = H[ix + 1, iy]
    h01 = H[ix, iy + 1]
    h11 = H[ix + 1, iy + 1]
    return (h00 * (1 - tx) * (1 - ty) + h10 * tx * (1 - ty) + h01 * (1 - tx) * ty + h11 * tx * ty)


def yaw_matrix(yaw):
    c, s = math.cos(yaw), math.sin(yaw)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def rot_xyz(rx, ry, rz):
    cx, sx, cy, sy = math.cos(rx), math.sin(rx), math.cos(ry), math.sin(ry)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    return yaw_matrix(rz) @ Ry @ Rx


def lidar_scan(world: World, raster, origin, R, rings, azimuths, elev_min_deg, elev_max_deg,
               r_min, r_max, dropout, rng, noise=0.02, step=0.5):
    """Returns world-frame f64 points (kept returns only), ring-major order."""
    elev = np.deg2rad(np.linspace(elev_min_deg, elev_max_deg, rings))
    az = rng.uniform(0, 2 * np.pi) + 2 * np.pi * np.arange(azimuths) / azimuths
    E, A = np.meshgrid(elev, az, indexing="ij")          # ring-major
    d_s = np.stack([np.cos(E) * np.cos(A), np.cos(E) * np.sin(A), np.sin(E)], -1).reshape(-1, 3)
    d = d_s @ R.T
    n = d.shape[0]
    o = np.asarray(origin, np.float64)
    hmax = float(raster[0].max()) + 0.5
    rng_hit = np.full(n, np.nan)
    active = np.arange(n)
    prev_r = np.full(n, r_min)
    r = r_min
    while active.size and r <= r_max:
        p = o + r * d[active]
        h = _sample(raster, p[:, 0], p[:, 1])
        hit = p[:, 2] <= h
        idx = active[hit]
        # bisection refine between prev_r and r
        lo = prev_r[idx].copy()
        hi = np.full(idx.size, r)
        for _ in range(12):
            mid = 0.5 * (lo + hi)
            q = o + mid[:, None] * d[idx]
            below = q[:, 2] <= _sample(raster, q[:, 0], q[:, 1])
            hi = np.where(below, mid, hi)
            lo = np.where(below, lo, mid)
        rng_hit[idx] = hi
        escaped = (p[:, 2] > hmax) & (d[active, 2] >= 0)
        keep = ~hit & ~escaped
        prev_r[active[keep]] = r
        active = active[keep]
        r += step
    # vegetation: exponential returns while inside 2 m-tall discs
    if world.veg.shape[0]:
        for cx, cy, rad in world.veg:
            # closest approach of the ray's xy line to the disc centre
            dxy = d[:, :2]
            nrm = np.maximum(np.linalg.norm(dxy, axis=1), 1e-9)
            u = dxy / nrm[:, None]
            oc = np.array([cx - o[0], cy - o[1]])
            tca = u @ oc
            d2 = oc @ oc - tca ** 2
            inside = d2 < rad * rad
            thc = np.sqrt(np.maximum(rad * rad - d2, 0))
            t0 = (tca - thc) / nrm
            t1 = (tca + thc) / nrm
            cand = inside & (t1 > r_min)
            if not cand.any():
                continue
            t_enter = np.maximum(t0, r_min)
            t_draw = t_enter + rng.exponential(1.0 / 0.3, n)
            ok = cand & (t_draw < t1) & (t_draw <= r_max)
            q = o + t_draw[:, None] * d
            hq = world.terrain(q[:, 0], q[:, 1])
            ok &= (q[:, 2] >= hq) & (q[:, 2] <= hq + 2.0)
            better = ok & (np.isnan(rng_hit) | (t_draw < rng_hit))
            rng_hit[better] = t_draw[better]
    far = np.isnan(rng_hit)
    rng_hit[far] = rng.uniform(0.5 * r_max, r_max, int(far.sum()))
    rng_hit = rng_hit + rng.normal(0.0, noise, n)
    rng_hit = np.clip(rng_hit, r_min, r_max)
    keep = rng.random(n) >= dropout
    pts = o + rng_hit[:, None] * d
    return pts[keep]


@dataclass
class ScanInput:
    points: np.ndarray       # (N, 3) f32 in the sensor frame (or world frame when T is identity)
    ego: np.ndarray          # (3,) f64 vehicle position (rays start here)
    transform: np.ndarray    # (4, 4) f32 sensor-to-world
    n_rays: int


@dataclass
class Workload:
    name: str
    map_args: tuple          # the 14 positional Gvom constructor arguments
    steps: list = field(default_factory=list)   # list of lists of ScanInput (one combine per step)


def gvom_args(xy_res, z_res, xy, z, buf):
    d = DEFAULTS
    return (xy_res, z_res, xy, z, buf, d["min_distance"], d["positive_obstacle_threshold"],
            d["negative_obstacle_threshold"], d["slope_obsacle_threshold"], d["robot_height"],
            d["robot_radius"], d["ground_to_lidar_height"], d["xy_eigen_dist"], d["z_eigen_dist"])


def _to_sensor(pts_world, Rw, tw):
    return ((pts_world - tw) @ Rw).astype(np.float32)


def _T(Rw, tw):
    T = np.eye(4)
    T[:3, :3] = Rw
    T[:3, 3] = tw
    return T.astype(np.float32)


def _scan_for(world, raster, rng, R_pose, t_pose, R_ext, t_ext, rings, az, emin, emax, rmin, rmax,
              dropout, ego):
    Rw = R_pose @ R_ext
    tw = R_pose @ t_ext + t_pose
    pw = lidar_scan(world, raster, tw, Rw, rings, az, emin, emax, rmin, rmax, dropout, rng)
    return ScanInput(_to_sensor(pw, Rw, tw), np.asarray(ego, np.float64), _T(Rw, tw), rings * az)


def workload(cfg: int, frames: int | None = None, scale_points: float = 1.0) -> Workload:
    """Configurations 1..5 of the scan->map benchmark, plus 0 = the headline
    (one 128x1024 sensor, 256x256x64 voxels, buffer 8, 8 frames at 4 m/s, 10 Hz)."""
    seeds = {0: 100, 1: 101, 2: 102, 3: 103, 4: 104, 5: 105}
    base = np.random.SeedSequence(seeds[cfg])
    pose_rng = np.random.default_rng(base.spawn(1)[0])

    def scan_rng(frame, sensor):
        return np.random.default_rng(np.random.SeedSequence([seeds[cfg], frame, sensor]))

    if cfg == 1:
        world = make_world(seeds[cfg], 2, 0, False, extent=30.0)
        raster = _raster(world, -35, 35, -35, 35)
        rng = scan_rng(0, 0)
        ego = np.array([0.0, 0.0, 1.0])
        pw = lidar_scan(world, raster, ego, np.eye(3), 16, 512, -15.0, 15.0, 1.0, 30.0, 0.05, rng)
        s = ScanInput(pw.astype(np.float32), ego, np.eye(4, dtype=np.float32), 16 * 512)
        return Workload("cfg1: 16x512 scan, 64x64x32 @0.4/0.2, buffer 1", gvom_args(0.4, 0.2, 64, 32, 1), [[s]])

    if cfg in (2, 5):
        world = make_world(seeds[2], 20, 5, False)
        nfr = 4 if cfg == 2 else 1
        nfr = frames or nfr
        raster = _raster(world, -105, 105 + nfr, -105, 105)
        wl = Workload(f"cfg{cfg}: 64x1024 scans, 256x256x64 @0.4/0.2, buffer 4", gvom_args(0.4, 0.2, 256, 64, 4))
        yaw = 0.0
        for f in range(nfr):
            yaw += math.radians(pose_rng.normal(0.0, 1.0))
            t = np.array([float(f), 0.0, 1.0 + float(world.terrain(float(f), 0.0))])
            s = _scan_for(world, raster, scan_rng(f, 0), yaw_matrix(yaw), t, np.eye(3), np.zeros(3),
                          64, 1024, -22.5, 22.5, 1.0, 100.0, 0.05, t)
            wl.steps.append([s])
        return wl

    if cfg == 3:
        world = make_world(seeds[3], 20, 5, True)
        nfr = frames or 8
        raster = _raster(world, -125, 130, -125, 130)
        wl = Workload("cfg3: 128x2048 scans, 512x512x64 @0.4/0.2, buffer 8", gvom_args(0.4, 0.2, 512, 64, 8))
        x = y = yaw = 0.0
        for f in range(nfr):
            t = np.array([x, y, 1.0 + float(world.terrain(x, y))])
            s = _scan_for(world, raster, scan_rng(f, 0), yaw_matrix(yaw), t, np.eye(3), np.zeros(3),
                          128, 2048, -22.5, 22.5, 1.0, 120.0, 0.10, t)
            wl.steps.append([s])
            x += 0.4 * math.cos(yaw)
            y += 0.4 * math.sin(yaw)
            yaw += 0.05
        return wl

    if cfg in (0, 4):
        world = make_world(seeds[cfg], 20, 5, False)
        nfr = frames or 8
        raster = _raster(world, -105, 105 + 0.4 * nfr, -105, 105)
        sensors = [(128, 1024, -22.5, 22.5), (64, 1024, -22.5, 22.5), (16, 1024, -15.0, 15.0)]
        if cfg == 0:
            sensors = sensors[:1]
        name = ("headline: 128x1024 scans, 256x256x64 @0.4/0.2, buffer 8" if cfg == 0 else
                "cfg4: 3 sensors (128/64/16x1024), 256x256x64 @0.4/0.2, buffer 8")
        wl = Workload(name, gvom_args(0.4, 0.2, 256, 64, 8))
        for f in range(nfr):
            x = 0.4 * f
            t = np.array([x, 0.0, 1.0 + float(world.terrain(x, 0.0))])
            Rp = yaw_matrix(0.0)
            step = []
            for k, (rings, az, emin, emax) in enumerate(sensors):
                yaw_off = math.radians(120.0 * k)
                t_ext = np.array([0.0, 0.5 * (k - 1), 0.0]) if cfg == 4 else np.zeros(3)
                dropout = 0.0 if cfg == 0 else float(pose_rng.uniform(0.10, 0.40))
                step.append(_scan_for(world, raster, scan_rng(f, k), Rp, t, yaw_matrix(yaw_off), t_ext,
                                      int(rings * scale_points) if scale_points != 1.0 else rings,
                                      az, emin, emax, 1.0, 100.0, dropout, t))
            wl.steps.append(step)
        return wl
    raise ValueError(f"unknown configuration {cfg}")


REFINE_POSES = ((0.0, 0.0, 0.0), (6.0, 3.0, 45.0), (12.0, 1.0, -30.0), (15.0, -5.0, -80.0))


def refine_frames(poses=REFINE_POSES, rings=64, azimuths=1024, seed=205):
    """Frames for extrinsic refinement on the configuration-5 world: 64 x 1024 scans taken
    with the true (identity) extrinsic at vehicle poses (x m, y m, heading deg) along an
    S-shaped path.  The headings differ, so a translation of the extrinsic moves each frame's
    points differently; the poses lie on no common circle, so no yaw of the extrinsic combined
    with a translation moves all frames as one rigid body (on a circular arc it would, and
    the loss could not tell that combination from the truth).
    Returns (sensor-frame points per frame (float64), vehicle poses (4x4 float64), ego
    positions)."""
    world = make_world(102, 20, 5, False)
    raster = _raster(world, -105, 125, -110, 110)
    pts, poses_out, egos = [], [], []
    for f, (x, y, hd) in enumerate(poses):
        th = math.radians(hd)
        t = np.array([x, y, 1.0 + float(world.terrain(x, y))])
        s = _scan_for(world, raster, np.random.default_rng(np.random.SeedSequence([seed, f])), yaw_matrix(th), t,
                      np.eye(3), np.zeros(3), rings, azimuths, -22.5, 22.5, 1.0, 100.0, 0.05, t)
        pts.append(s.points.astype(np.float64))
        poses_out.append(s.transform.astype(np.float64))
        egos.append(t)
    return pts, poses_out, egos


def perturb_transform(T, rng, rot_deg=0.5, trans_m=0.05):
    """Configuration 5: rotate by N(0, rot_deg) per axis and translate by N(0, trans_m)."""
    r = np.deg2rad(rng.normal(0.0, rot_deg, 3))
    dT = np.eye(4)
    dT[:3, :3] = rot_xyz(*r)
    dT[:3, 3] = rng.normal(0.0, trans_m, 3)
    return (dT @ T.astype(np.float64)).astype(np.float32)
