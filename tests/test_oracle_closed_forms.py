"""Closed-form pins of six oracle functions -- CPU only.

Each test fixes the oracle's answer from what the mathematics of the reference's rule
forces, not from a second restatement of the same formula, and each runs again with a
planted mistake in the oracle (oracle.gvom_oracle.planted_mistake) that must make it fail:

  - point transform (gvom.py:1122-1138): a translation by whole voxels of the points and
    the ego together moves the origin by those voxels and leaves the scan's occupancy,
    hits, minimum heights and metrics bit-identical; a 90 degree yaw about the world z axis
    through the map centre permutes them as numpy's rot90;
  - combine origin shift (gvom.py:1017-1026, 1044-1053): one scan combined after an empty
    newer scan moved the origin by d voxels equals the same combine without the move,
    shifted by d with unknown fill; combined once more after a further move, the history
    term doubles hits, totals and passes exactly where both shifts keep the voxel;
  - inferred height (gvom.py:580-590): a vertical ray from a voxel-centred ego makes the
    lowest free voxel of its column the one just above the hit;
  - robot-radius default height (gvom.py:566-570): with the ego on a cell corner and a
    radius of k cells, the cells given the default height are the lattice points of a disc
    (Gauss circle counts 5, 13, 29, 49, 81, 113);
  - debug exporters (gvom.py:455-503): rows hand-computed from a three-cell map whose
    hits, totals and eigenvalues are worked out by hand below;
  - the minimum-distance filter (gvom.py:1145-1149): measured from the world origin after the
    transform, not from the ego, with |p| = min_distance kept.

All coordinates are dyadic, so every float in these scenes is exact and the expected
values are exact.
"""
import numpy as np
import pytest

from oracle.gvom_oracle import OracleGvom, canonical_scan, planted_mistake

XYR, ZR = 0.5, 0.25


def _params(xy, z, B, robot_radius=0.0, min_distance=0.0):
    # (xy_res, z_res, xy, z, B, min_distance, pos, neg, slope, robot_h, robot_r, g2l, ex, ez)
    return (XYR, ZR, xy, z, B, min_distance, 0.5, 0.5, 0.3, 2.0, robot_radius, 1.0, 1, 1)


def _dense(o, slot, xy, z):
    """Per-voxel fields of a scan slot as [x, y, z] arrays (reference layout x + y*xy + z*xy^2)."""
    c = canonical_scan(slot.index_map, slot.hit, slot.total, slot.min_h, slot.metrics)

    def cube(a):
        return a.reshape((z, xy, xy) + a.shape[1:]).transpose((2, 1, 0) + tuple(range(3, a.ndim + 2)))
    return {k: cube(v) for k, v in c.items() if k != "cells"}


def _dyadic_scan(rng, n, lo, hi, zlo, zhi):
    """Points with coordinates on a 1/16-voxel lattice strictly inside voxels."""
    i = rng.integers(lo, hi, n)
    j = rng.integers(lo, hi, n)
    k = rng.integers(zlo, zhi, n)
    fx, fy, fz = (rng.integers(1, 16, (3, n)) / 16.0)
    return np.stack([(i + fx) * XYR, (j + fy) * XYR, (k + fz) * ZR], axis=1)


# -- point transform -------------------------------------------------------------------
def _translation_case(dtype, planted):
    xy, z = 32, 16
    rng = np.random.default_rng(7)
    pts = _dyadic_scan(rng, 400, -10, 10, -4, 4).astype(dtype)
    ego = np.array([0.75, -1.25, 0.375])
    a, b, c = 3, -2, 1
    T = np.eye(4, dtype=dtype)
    T[:3, 3] = (a * XYR, b * XYR, c * ZR)
    o1 = OracleGvom(*_params(xy, z, 1))
    o2 = OracleGvom(*_params(xy, z, 1))
    with planted_mistake(planted):
        o1.Process_pointcloud(pts, ego, None)
        o2.Process_pointcloud(pts, ego + T[:3, 3].astype(np.float64), T)
    s1, s2 = o1.slots[0], o2.slots[0]
    d1, d2 = _dense(o1, s1, xy, z), _dense(o2, s2, xy, z)
    ok = np.array_equal(s2.origin, s1.origin + (a, b, c))
    for k in ("occ", "hit", "min_h", "metrics"):
        ok = ok and np.array_equal(d1[k], d2[k], equal_nan=True)
    return ok, int(d1["occ"].sum())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_transform_translation_by_whole_voxels(dtype):
    ok, cells = _translation_case(dtype, 0)
    assert cells > 50
    assert ok, "translating points and ego by whole voxels must only move the origin"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_transform_translation_pin_catches_a_dropped_translation(dtype):
    ok, _ = _translation_case(dtype, 5)
    assert not ok


def _yaw_case(dtype, planted):
    xy, z = 32, 16
    rng = np.random.default_rng(9)
    pts = _dyadic_scan(rng, 400, -12, 12, -4, 4).astype(dtype)
    ego = np.array([0.0, 0.0, 0.375])          # origin xy = -xy/2: the map centre is the world z axis
    R = np.eye(4, dtype=dtype)
    R[:2, :2] = [[0, -1], [1, 0]]              # (x, y) -> (-y, x)
    o1 = OracleGvom(*_params(xy, z, 1))
    o2 = OracleGvom(*_params(xy, z, 1))
    with planted_mistake(planted):
        o1.Process_pointcloud(pts, ego, None)
        o2.Process_pointcloud(pts, ego, R)
    d1, d2 = _dense(o1, o1.slots[0], xy, z), _dense(o2, o2.slots[0], xy, z)

    def rot(a):
        return np.rot90(a, 1, axes=(0, 1))
    ok = all(np.array_equal(rot(d1[k]), d2[k], equal_nan=True) for k in ("occ", "hit", "min_h"))
    if ok:
        # metrics columns: 0-2 mean, 3-8 covariance xx xy xz yy yz zz, 9 count (gvom.py:1355-1371).
        # A rotated point's local coordinates are (1 - l_y, l_x, l_z).
        m1, m2 = rot(d1["metrics"]), d2["metrics"]
        occ = d2["occ"]
        a, b = m1[occ], m2[occ]
        expect = np.stack([1 - a[:, 1], a[:, 0], a[:, 2], a[:, 6], -a[:, 4], -a[:, 7], a[:, 3], a[:, 5], a[:, 8],
                           a[:, 9]], axis=1)
        ok = np.allclose(b, expect, rtol=0, atol=1e-12)
    return ok, int(d1["occ"].sum())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_transform_yaw_90_permutes_the_table(dtype):
    ok, cells = _yaw_case(dtype, 0)
    assert cells > 50
    assert ok


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_transform_yaw_pin_catches_a_transposed_rotation(dtype):
    ok, _ = _yaw_case(dtype, 1)
    assert not ok


# -- combine origin shift -------------------------------------------------------------
def _combined_dense(o, xy, z):
    cm = o.combined
    c = canonical_scan(cm.index_map, cm.hit, cm.total, cm.min_h, cm.metrics)
    ev = np.full((cm.index_map.shape[0], 3), np.nan, np.float32)
    occ = cm.index_map >= 0
    ev[occ] = cm.eigen[cm.index_map[occ]]
    c["eigen"] = ev

    def cube(a):
        return a.reshape((z, xy, xy) + a.shape[1:]).transpose((2, 1, 0) + tuple(range(3, a.ndim + 2)))
    return {k: cube(v) for k, v in c.items() if k != "cells"}


def _shift(a, d, fill):
    """out[x] = a[x + d] where x + d is inside the map, else fill (per axis of x, y, z)."""
    out = np.full_like(a, fill)
    src = [slice(max(0, dk), a.shape[k] + min(0, dk)) for k, dk in enumerate(d)]
    dst = [slice(max(0, -dk), a.shape[k] + min(0, -dk)) for k, dk in enumerate(d)]
    out[tuple(dst)] = a[tuple(src)]
    return out


def _inside(shape, d):
    m = np.zeros(shape[:3], bool)
    m[tuple(slice(max(0, -dk), shape[k] + min(0, -dk)) for k, dk in enumerate(d))] = True
    return m


FILL = {"occ": False, "passes": 0, "hit": 0, "total": 0, "min_h": np.nan, "metrics": np.nan, "eigen": np.nan}


def _shift_case(planted):
    xy, z = 24, 16
    rng = np.random.default_rng(3)
    pts = _dyadic_scan(rng, 600, -9, 9, -4, 4)
    ego = np.array([0.25, 0.25, 0.125])
    d = (3, -2, 1)
    d2 = (-1, 4, -2)
    step = np.array([XYR, XYR, ZR])
    empty = np.zeros((0, 3))
    base = OracleGvom(*_params(xy, z, 2))
    moved = OracleGvom(*_params(xy, z, 3))
    with planted_mistake(planted):
        base.Process_pointcloud(pts, ego)
        base.Process_pointcloud(empty, ego)
        base.combine_maps()
        moved.Process_pointcloud(pts, ego)
        moved.Process_pointcloud(empty, ego + np.multiply(d, step))
        moved.combine_maps()
        c0 = _combined_dense(base, xy, z)
        c1 = _combined_dense(moved, xy, z)
        moved.Process_pointcloud(empty, ego + np.multiply(np.add(d, d2), step))
        moved.combine_maps()
        c2 = _combined_dense(moved, xy, z)
    ok1 = all(np.array_equal(_shift(c0[k], d, FILL[k]), c1[k], equal_nan=True) for k in FILL)
    # second combine: the scan shifted by d + d2, plus the previous combined map shifted by d2
    # (itself the scan shifted by d) -- twice the scan's counts where x + d2 is inside the map
    dd = tuple(np.add(d, d2))
    twice = np.where(_inside(c0["hit"].shape, d2), 2, 1)
    ok2 = np.array_equal(_shift(c0["occ"], dd, False), c2["occ"])
    for k in ("hit", "total", "passes"):
        ok2 = ok2 and np.array_equal(_shift(c0[k], dd, 0) * twice, c2[k])
    ok2 = ok2 and np.array_equal(_shift(c0["min_h"], dd, np.nan), c2["min_h"], equal_nan=True)
    return ok1, ok2, int(c0["occ"].sum()), int((c0["passes"] > 0).sum())


def test_combine_shift_equals_shifted_table():
    ok1, ok2, cells, free = _shift_case(0)
    assert cells > 50 and free > 50
    assert ok1, "combine after an origin move must equal the unmoved combine shifted by d"
    assert ok2, "the history term must double counts where both shifts keep the voxel"


def test_combine_shift_pin_catches_a_reversed_shift():
    ok1, ok2, _, _ = _shift_case(2)
    assert not ok1 and not ok2


# -- inferred height ------------------------------------------------------------------
def _vertical_case(a, hits, planted):
    """ego at the centre of voxel height a (world z = (a + 0.5) * ZR), points straight below
    at voxel heights `hits` (centres).  Returns the oracle's height / inferred maps and the
    ego column."""
    xy, z = 16, 32
    o = OracleGvom(*_params(xy, z, 1))
    ego = np.array([1.25, -0.75, (a + 0.5) * ZR])          # xy at a cell centre
    pts = np.array([[ego[0], ego[1], (b + 0.5) * ZR] for b in hits])
    with planted_mistake(planted):
        o.Process_pointcloud(pts, ego)
        o.combine_maps()
    oz = o.slots[0].origin[2]
    col = (int(np.floor(ego[0] / XYR - o.slots[0].origin[0])), int(np.floor(ego[1] / XYR - o.slots[0].origin[1])))
    return o.combined, col, oz


@pytest.mark.parametrize("a,hits,lowest_free", [
    (10, [0], 1),            # 10 voxels down: rays visit 9 .. 1
    (10, [8], 9),
    (10, [9], None),         # one voxel down: ceil(L - 1) = 0 steps, no pass
    (12, [2, 6], 3),         # the upper hit lies on the lower point's ray; lowest free = 3
    (12, [2, 3], 4),         # the voxel just above the lower hit is occupied: lowest free = 4
])
def test_inferred_height_of_a_vertical_ray(a, hits, lowest_free):
    cm, col, oz = _vertical_case(a, hits, 0)
    expect = -1000.0 if lowest_free is None else lowest_free * ZR
    assert cm.inferred[col] == expect
    assert cm.height[col] == (min(hits) + 0.5) * ZR     # lowest point of the lowest occupied voxel
    others = np.ones(cm.inferred.shape, bool)
    others[col] = False
    assert (cm.inferred[others] == -1000.0).all()


def test_inferred_height_pin_catches_an_off_by_one():
    cm, col, _ = _vertical_case(10, [0], 3)
    assert cm.inferred[col] != 1 * ZR


# -- robot-radius default height ------------------------------------------------------
GAUSS_CIRCLE = {1: 5, 2: 13, 3: 29, 4: 49, 5: 81, 6: 113}   # lattice points with i^2 + j^2 <= k^2


def _radius_case(k, planted):
    xy, z = 32, 16
    o = OracleGvom(*_params(xy, z, 1, robot_radius=k * XYR))
    ego = np.array([1.5, -2.0, 0.625])                   # on a cell corner (xy multiples of XYR)
    with planted_mistake(planted):
        o.Process_pointcloud(np.zeros((0, 3)), ego)
        out = o.combine_maps()
    return o.combined, out, ego, o.slots[0].origin


@pytest.mark.parametrize("k", sorted(GAUSS_CIRCLE))
def test_robot_radius_cells_are_the_lattice_disc(k):
    cm, out, ego, origin = _radius_case(k, 0)
    xy = cm.height.shape[0]
    cx, cy = int(ego[0] / XYR - origin[0]), int(ego[1] / XYR - origin[1])   # the cell whose corner is ego
    i, j = np.meshgrid(np.arange(xy) - cx, np.arange(xy) - cy, indexing="ij")
    disc = i * i + j * j <= k * k
    assert disc.sum() == GAUSS_CIRCLE[k]
    np.testing.assert_array_equal(cm.height > -1000.0, disc)
    assert (cm.height[disc] == ego[2] - 1.0).all()           # ego z - ground_to_lidar_height
    np.testing.assert_array_equal(out[4], disc.astype(np.int32))   # visibility


def test_robot_radius_pin_catches_the_cell_centre():
    cm, _, _, _ = _radius_case(1, 4)
    assert (cm.height > -1000.0).sum() != GAUSS_CIRCLE[1]


# -- minimum-distance filter -----------------------------------------------------------
def _min_distance_case(planted):
    """min_distance 2 m, ego at (6, 0, 0.5): points within 2 m of the world origin are dropped
    (no hit, no ray) although they are in the map and far from the ego; points within 2 m of
    the ego are kept (gvom.py:1145-1149 measures from the world origin after the transform)."""
    xy, z = 32, 16
    o = OracleGvom(*_params(xy, z, 1, min_distance=2.0))
    ego = np.array([6.0, 0.0, 0.5])
    near_origin = np.array([[0.5625, 0.5625, 0.3125], [-1.0625, 0.8125, 0.3125], [1.8125, 0.0625, 0.0625]])
    on_circle = np.array([[2.0, 0.0, 0.0]])                # |p| = 2 exactly: d2 < 4 is false, kept
    near_ego = ego + np.array([[0.3125, 0.3125, 0.0625], [-0.5625, 0.8125, -0.1875]])
    pts = np.concatenate([near_origin, on_circle, near_ego])
    with planted_mistake(planted):
        o.Process_pointcloud(pts, ego)
    sl = o.slots[0]
    occ = canonical_scan(sl.index_map, sl.hit, sl.total)["occ"]
    org = sl.origin

    def vox(p):
        f = np.floor(np.array([p[0] / XYR, p[1] / XYR, p[2] / ZR]) - org).astype(int)
        return f[0] + f[1] * xy + f[2] * xy * xy
    kept = [bool(occ[vox(p)]) for p in pts]
    return kept, int(occ.sum()), int(sl.valid.sum())


def test_min_distance_is_measured_from_the_world_origin():
    kept, cells, valid = _min_distance_case(0)
    assert kept == [False, False, False, True, True, True]
    assert valid == 3 and cells == 3


def test_min_distance_pin_catches_a_filter_around_the_ego():
    kept, _, _ = _min_distance_case(9)
    assert kept != [False, False, False, True, True, True]


# -- debug exporters ------------------------------------------------------------------
def _three_cells(planted):
    """Voxel A = (8, 8, 8) with its corner at world (2.0, -3.0, 0.5); ego at A's centre.
    Points in A-local voxel units (x, y, z):
      A: (0.25, 0.5, 0.5), (0.75, 0.5, 0.5)   -- within a quarter voxel of ego: no ray step;
      B = A + x: (1.5, 0.25, 0.5), (1.5, 0.75, 0.5) -- ray length 1.03: one step, into B;
      C = A + 3x: (3.5, 0.5, 0.5)              -- ray length 3: two steps, into B and A + 2x.
    So hits / totals are A 2/2, B 2/5, C 1/1.  A's and B's neighbourhoods hold the four
    points of A and B (C is two voxels from B): mean (1, 0.5, 0.5), variances 0.28125 (x)
    and 0.03125 (y), no covariance, so eigenvalues (0.28125, 0.03125, 0); C's holds one
    point: (0, 0, 0).  Height of A, B, C: their points' z, 0.625; windows hold < 3 observed
    columns, so roughness -1 and slopes 0 everywhere."""
    xy, z = 16, 16
    o = OracleGvom(*_params(xy, z, 1))
    corner = np.array([2.0, -3.0, 0.5])
    unit = np.array([XYR, XYR, ZR])
    local = np.array([[0.25, 0.5, 0.5], [0.75, 0.5, 0.5], [1.5, 0.25, 0.5], [1.5, 0.75, 0.5], [3.5, 0.5, 0.5]])
    ego = corner + 0.5 * unit
    with planted_mistake(planted):
        o.Process_pointcloud(corner + local * unit, ego)
        o.combine_maps()
        vox = o.make_debug_voxel_map()
        hgt = o.make_debug_height_map()
        inf = o.make_debug_inferred_height_map()
    return o, vox, hgt, inf


def _rows_ok(o, vox, hgt, inf):
    cm = o.combined
    xy = cm.height.shape[0]
    lin = lambda x, y, z: x + y * xy + z * xy * xy   # noqa: E731
    want = {
        (8, 8, 8): [2.0, -3.0, 0.5, 1.0, 2, 0.25, 0.03125, 0.0],
        (9, 8, 8): [2.5, -3.0, 0.5, np.float32(2 / 5), 2, 0.25, 0.03125, 0.0],
        (11, 8, 8): [3.5, -3.0, 0.5, 1.0, 1, 0.0, 0.0, 0.0],
    }
    if vox.shape != (3, 8):
        return False
    for v, row in want.items():
        if not np.array_equal(vox[cm.index_map[lin(*v)]], np.array(row, np.float32)):
            return False
    # height-map rows: index x + y * xy; (x, y) = B's column and its transpose (8, 9)
    hrows = {9 + 8 * xy: [2.5, -3.0, 0.375, -1.0, 0.0, 0.0, 0.0],
             8 + 9 * xy: [2.0, -2.5, -1000.25, -1.0, 0.0, 0.0, 0.0]}
    for r, row in hrows.items():
        if not np.array_equal(hgt[r], np.array(row, np.float32)):
            return False
    # inferred rows: the guessed-height delta (0 where no guess is made) minus z_res
    irows = {9 + 8 * xy: [2.5, -3.0, -0.25], 8 + 9 * xy: [2.0, -2.5, -0.25]}
    return all(np.array_equal(inf[r], np.array(row, np.float32)) for r, row in irows.items())


def test_debug_exporter_rows_hand_computed():
    o, vox, hgt, inf = _three_cells(0)
    assert _rows_ok(o, vox, hgt, inf)


@pytest.mark.parametrize("planted", [6, 7, 8])
def test_debug_exporter_pin_catches_planted_mistakes(planted):
    o, vox, hgt, inf = _three_cells(planted)
    assert not _rows_ok(o, vox, hgt, inf)
