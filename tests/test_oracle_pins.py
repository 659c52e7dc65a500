"""Pins the oracle (oracle/gvom_oracle.c) -- CPU only.

The reference cannot run here (its only numba install fails to import), so the
oracle is checked against what fixes its answers independently:
  - golden vectors from a second, independent restatement (oracle/pyref.py),
    compared bit-exactly;
  - closed forms: np.bincount hit counts, np.minimum.at minimum heights, two-pass
    numpy neighbourhood mean / population covariance, LAPACK eigenvalues, the pooled
    (Chan) covariance of two point sets, axis-aligned ray step counts, the plane-fit
    slope and roughness of a tilted plane;
  - hand-derived answers for the reference's quirks (carry-over rule, history
    double counting, the __guess_height loop condition, scan ranges and y_nh gating).
"""
import glob
import math
import os

import numpy as np
import pytest

from oracle import gvom_oracle as go
from oracle.gvom_oracle import OracleGvom, canonical_scan
from oracle.pyref import PyRef

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# -- golden vectors (independent restatement) ---------------------------------------------
def _frames(z):
    for f in range(int(z["frames"])):
        T = z[f"f{f}_T"]
        yield f, z[f"f{f}_points"], z[f"f{f}_ego"], (T if T.size else None)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_oracle_matches_golden_vectors(path):
    z = np.load(path)
    o = OracleGvom(*[float(v) if i in (0, 1, 5, 6, 7, 8, 9, 10, 11) else int(v)
                     for i, v in enumerate(z["params"])])
    for f, pts, ego, T in _frames(z):
        o.Process_pointcloud(pts, ego, T)
        s = o.slots[o.last_buffer_index]
        for k in ("index_map", "hit", "total", "min_h", "origin"):
            np.testing.assert_array_equal(getattr(s, k), z[f"f{f}_slot_{k}"], err_msg=f"frame {f} slot {k}")
        np.testing.assert_array_equal(s.metrics, z[f"f{f}_slot_metrics"], err_msg=f"frame {f} metrics")
        o.combine_maps()
        c = o.combined
        for k in ("index_map", "hit", "total", "min_h", "metrics", "eigen", "origin", "height", "inferred",
                  "slope_x", "slope_y", "rough", "delta", "pos", "neg", "vis"):
            np.testing.assert_array_equal(getattr(c, k), z[f"f{f}_comb_{k}"], err_msg=f"frame {f} combined {k}")


def test_pyref_and_oracle_agree_on_random_tiny_scenes():
    rng = np.random.default_rng(1234)
    for trial in range(3):
        params = (0.5, 0.25, 10, 8, 2, 0.7, 0.5, 0.5, 0.3, 2.0, 1.2, 1.0, 1, 1)
        o = OracleGvom(*params)
        r = PyRef(*params)
        for f in range(3):
            ego = rng.uniform(-1, 1, 3)
            pts = (rng.normal(0, 1.5, (150, 3)) + ego).astype(np.float32)
            pts[:, 2] = rng.uniform(-1.5, 1.5, 150)
            o.Process_pointcloud(pts, ego)
            r.process(pts, ego)
            so, sr = o.slots[o.last_buffer_index], r.slots[r.last_buffer_index]
            np.testing.assert_array_equal(so.index_map, sr["index_map"])
            np.testing.assert_array_equal(so.total, sr["total"])
            np.testing.assert_array_equal(so.metrics, sr["metrics"])
            o.combine_maps()
            res = r.combine()
            for k in ("index_map", "hit", "total", "metrics", "eigen", "height", "rough", "delta", "pos", "neg"):
                np.testing.assert_array_equal(getattr(o.combined, k), res[k], err_msg=f"trial {trial} frame {f} {k}")


# -- closed forms on a realistic scan ------------------------------------------------------
@pytest.fixture(scope="module")
def cfg1_scan():
    import gvom_synth
    wl = gvom_synth.workload(1)
    s = wl.steps[0][0]
    o = OracleGvom(*wl.map_args)
    o.Process_pointcloud(s.points, s.ego, s.transform)
    return wl, s, o


def _own_voxels(o, slot):
    p = o.p
    w = slot.points_world.astype(np.float64)
    q = np.column_stack([w[:, 0] / p.xy_resolution - slot.origin[0], w[:, 1] / p.xy_resolution - slot.origin[1],
                         w[:, 2] / p.z_resolution - slot.origin[2]])
    return q, np.floor(q)


def test_hit_counts_equal_bincount(cfg1_scan):
    wl, s, o = cfg1_scan
    slot = o.slots[0]
    q, f = _own_voxels(o, slot)
    XY, Z = o.xy_size, o.z_size
    inb = slot.valid & (f[:, 0] >= 0) & (f[:, 0] < XY) & (f[:, 1] >= 0) & (f[:, 1] < XY) & (f[:, 2] >= 0) & (f[:, 2] < Z)
    fi = f[inb].astype(np.int64)
    v = fi[:, 0] + fi[:, 1] * XY + fi[:, 2] * XY * XY
    expect = np.bincount(v, minlength=o.voxel_count)
    c = canonical_scan(slot.index_map, slot.hit, slot.total)
    np.testing.assert_array_equal(c["hit"], expect)
    np.testing.assert_array_equal(c["occ"], expect > 0)
    assert c["cells"] == len(np.unique(v))


def test_min_height_equals_minimum_at(cfg1_scan):
    wl, s, o = cfg1_scan
    slot = o.slots[0]
    q, f = _own_voxels(o, slot)
    XY, Z = o.xy_size, o.z_size
    inb = slot.valid & (f[:, 0] >= 0) & (f[:, 0] < XY) & (f[:, 1] >= 0) & (f[:, 1] < XY) & (f[:, 2] >= 0) & (f[:, 2] < Z)
    fi = f[inb].astype(np.int64)
    v = fi[:, 0] + fi[:, 1] * XY + fi[:, 2] * XY * XY
    mh = np.ones(o.voxel_count, np.float32)
    np.minimum.at(mh, v, (q[inb, 2] - f[inb, 2]).astype(np.float32))
    c = canonical_scan(slot.index_map, slot.hit, slot.total, slot.min_h)
    occ = c["occ"]
    np.testing.assert_array_equal(c["min_h"][occ], mh[occ])


def test_metrics_equal_two_pass_numpy(cfg1_scan):
    """Mean and population covariance over the +-1 voxel neighbourhood, per occupied voxel."""
    wl, s, o = cfg1_scan
    slot = o.slots[0]
    q, f = _own_voxels(o, slot)
    XY, Z = o.xy_size, o.z_size
    c = canonical_scan(slot.index_map, slot.hit, slot.total, None, slot.metrics)
    occ_v = np.nonzero(c["occ"])[0]
    ok = slot.valid
    qv, fv = q[ok], f[ok]
    checked = 0
    for v in occ_v[::7]:   # a sample of voxels keeps this fast
        ux, uy, uz = v % XY, (v // XY) % XY, v // (XY * XY)
        sel = ((np.abs(fv[:, 0] - ux) <= 1) & (np.abs(fv[:, 1] - uy) <= 1) & (np.abs(fv[:, 2] - uz) <= 1))
        loc = qv[sel] - np.array([ux, uy, uz])
        mean = loc.mean(axis=0)
        cov = np.cov(loc.T, bias=True).reshape(3, 3) if len(loc) > 1 else np.zeros((3, 3))
        got = c["metrics"][v]
        np.testing.assert_allclose(got[:3], mean, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(got[[3, 4, 5, 6, 7, 8]], cov[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]],
                                   rtol=1e-8, atol=1e-12)
        assert got[9] == len(loc)
        checked += 1
    assert checked > 50


def test_ray_passes_total_equals_step_traces(cfg1_scan):
    """Sum of total counts minus hits = number of ray steps (traced by the independent
    restatement), and every ray's voxels are 26-connected with the dominant axis moving
    exactly one voxel per step."""
    wl, s, o = cfg1_scan
    slot = o.slots[0]
    r = PyRef(*wl.map_args)
    XY, Z = o.xy_size, o.z_size
    steps = 0
    rng = np.random.default_rng(0)
    sample = rng.choice(np.nonzero(slot.valid)[0], 400, replace=False)
    for i in sample:
        vox = r.ray_voxels([np.float32(v) for v in slot.points_world[i]], tuple(s.ego), tuple(slot.origin))
        xyz = np.array([(v % XY, (v // XY) % XY, v // (XY * XY)) for v in vox]).reshape(-1, 3)
        if len(xyz) > 1:
            d = np.abs(np.diff(xyz, axis=0))
            assert d.max() <= 1
    c = canonical_scan(slot.index_map, slot.hit, slot.total)
    total_steps = int(c["total"].sum() - c["hit"].sum() + c["passes"].sum())
    for i in np.nonzero(slot.valid)[0]:
        steps += len(r.ray_voxels([np.float32(v) for v in slot.points_world[i]], tuple(s.ego), tuple(slot.origin)))
    assert total_steps == steps


def _single_ray_map(points, ego, xy=32, z=16, res=(0.5, 0.25)):
    o = OracleGvom(res[0], res[1], xy, z, 1, 0.1, 0.5, 0.5, 0.3, 2.0, 0.5, 1.0, 1, 1)
    o.Process_pointcloud(np.asarray(points, np.float32), ego)
    return o


@pytest.mark.parametrize("k", [1, 2, 3, 7, 12])
def test_axis_aligned_ray_takes_ceil_L_minus_1_steps(k):
    """A ray of voxel length L along +x takes ceil(L - 1) steps (gvom.py:1208, 1231) and
    visits the voxels strictly between the ego voxel and the hit voxel."""
    ego = (0.25, 0.25, 0.125)          # exactly 0.5 voxel units on every axis
    p = (0.25 + 0.5 * k, 0.25, 0.125)  # L = k voxels
    o = _single_ray_map([p], ego)
    slot = o.slots[0]
    c = canonical_scan(slot.index_map, slot.hit, slot.total)
    XY = 32
    ox = int(slot.origin[0])
    ex = int(math.floor(0.5 - ox))
    counts = {}
    for v in np.nonzero(c["total"] + c["passes"])[0]:
        counts[(v % XY, (v // XY) % XY, v // (XY * XY))] = int(c["total"][v] + c["passes"][v])
    ey = int(math.floor(0.5 - slot.origin[1]))
    ez = int(math.floor(0.5 - slot.origin[2]))
    expect = {(ex + j, ey, ez): 1 for j in range(1, max(0, math.ceil(k - 1)) + 1)}
    hit = (ex + k, ey, ez)
    expect[hit] = expect.get(hit, 0) + 1
    assert counts == expect


@pytest.mark.parametrize("p,dom", [((0.25 + 2.0, 0.25 + 2.0, 0.125), 1), ((0.25 + 2.0, 0.25 + 2.0, 0.125 + 1.0), 2),
                                   ((0.25 - 2.0, 0.25 + 1.0, 0.125), 0)])
def test_dominant_axis_tie_break(p, dom):
    """Ties in |slope| go to y over x and to z over both (gvom.py:1199-1203): every step
    moves the dominant coordinate by exactly one voxel."""
    r = PyRef(0.5, 0.25, 32, 16, 1, 0.1, 0.5, 0.5, 0.3, 2.0, 0.5, 1.0, 1, 1)
    ego = (0.25, 0.25, 0.125)
    origin = (np.floor(0.5 - 16), np.floor(0.5 - 16), np.floor(0.5 - 8))
    vox = r.ray_voxels([np.float32(v) for v in p], ego, origin)
    xyz = np.array([(v % 32, (v // 32) % 32, v // 1024) for v in vox])
    assert len(xyz) >= 2
    d = np.diff(xyz, axis=0)
    assert np.all(np.abs(d[:, dom]) == 1)
    # the oracle counts the same voxels
    o = _single_ray_map([p], ego)
    c = canonical_scan(o.slots[0].index_map, o.slots[0].hit, o.slots[0].total)
    for v in vox:
        assert c["total"][v] + c["passes"][v] >= 1


# -- eigenvalues --------------------------------------------------------------------------
def test_eigenvalues_match_lapack():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(500, 3, 3))
    cov = (A @ np.transpose(A, (0, 2, 1))) * rng.uniform(0.001, 0.2, (500, 1, 1))
    m = np.zeros((500, 10), np.float32)
    m[:, 3:9] = cov[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    ev = go.eigenvalues(m)
    c32 = m[:, 3:9].astype(np.float64)
    full = np.zeros((500, 3, 3))
    for k, (a, b) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))):
        full[:, a, b] = c32[:, k]
        full[:, b, a] = c32[:, k]
    ref = np.linalg.eigvalsh(full)[:, ::-1]
    tr = np.trace(full, axis1=1, axis2=2)[:, None]
    assert np.all(np.abs(ev - ref) < 1e-5 * tr + 1e-7)


def test_eigenvalue_special_cases():
    m = np.zeros((4, 10), np.float32)
    m[0, [3, 6, 8]] = [0.2, 0.5, 0.1]                 # diagonal: sorted diagonal, exactly
    m[1, [3, 6, 8]] = [0.3, 0.3, 0.3]                 # isotropic
    v = np.array([0.3, -0.2, 0.5])
    vv = np.outer(v, v)                               # rank one: |v|^2, 0, 0
    m[2, 3:9] = vv[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    m[3, 3:9] = 0                                     # single point: all zero
    ev = go.eigenvalues(m)
    # extremes are the diagonal exactly; the middle one is 3q - max - min (gvom.py:1452-1456)
    assert ev[0, 0] == np.float32(0.5) and ev[0, 2] == np.float32(0.1)
    np.testing.assert_allclose(ev[0, 1], 0.2, rtol=2e-7)
    np.testing.assert_allclose(ev[1], [0.3, 0.3, 0.3], rtol=1e-6)
    np.testing.assert_allclose(ev[2], [v @ v, 0, 0], atol=2e-7)
    np.testing.assert_array_equal(ev[3], 0)


# -- combine -------------------------------------------------------------------------------
def test_single_slot_combine_is_the_scan(cfg1_scan):
    wl, s, o = cfg1_scan
    o2 = OracleGvom(*wl.map_args)
    o2.Process_pointcloud(s.points, s.ego, s.transform)
    o2.combine_maps()
    sl, cm = o2.slots[0], o2.combined
    a = canonical_scan(sl.index_map, sl.hit, sl.total, sl.min_h, sl.metrics)
    b = canonical_scan(cm.index_map, cm.hit, cm.total, cm.min_h, cm.metrics)
    for k in ("occ", "passes", "hit", "total"):
        np.testing.assert_array_equal(a[k], b[k])
    occ = a["occ"]
    np.testing.assert_array_equal(a["min_h"][occ], b["min_h"][occ])
    np.testing.assert_allclose(b["metrics"][occ], a["metrics"][occ], rtol=1e-5, atol=1e-7)


def test_history_is_merged_again_each_combine(cfg1_scan):
    """gvom.py:253-257, 295-298: the previous combined map is merged on every combine,
    so a second combine with no new scan doubles hits, totals and free passes."""
    wl, s, o = cfg1_scan
    o2 = OracleGvom(*wl.map_args)
    o2.Process_pointcloud(s.points, s.ego, s.transform)
    o2.combine_maps()
    first = canonical_scan(o2.combined.index_map, o2.combined.hit, o2.combined.total)
    o2.combine_maps()
    second = canonical_scan(o2.combined.index_map, o2.combined.hit, o2.combined.total)
    np.testing.assert_array_equal(second["occ"], first["occ"])
    np.testing.assert_array_equal(second["hit"], 2 * first["hit"])
    np.testing.assert_array_equal(second["total"], 2 * first["total"])
    np.testing.assert_array_equal(second["passes"], 2 * first["passes"])


def test_pooled_covariance_of_two_scans():
    """Merging two scans of the same voxel gives the mean and population covariance of
    the union of their neighbourhood point sets (gvom.py:921-965)."""
    rng = np.random.default_rng(3)
    params = (1.0, 1.0, 16, 8, 2, 0.5, 0.5, 0.5, 0.3, 2.0, 0.5, 1.0, 0, 0)   # eigen radius 0: own voxel only
    o = OracleGvom(*params)
    ego = np.array([0.5, 0.5, 0.5])
    a = (np.array([3.0, 2.0, 0.0]) + rng.uniform(0.05, 0.95, (30, 3))).astype(np.float32)
    b = (np.array([3.0, 2.0, 0.0]) + rng.uniform(0.05, 0.95, (45, 3))).astype(np.float32)
    o.Process_pointcloud(a, ego)
    o.Process_pointcloud(b, ego)
    o.combine_maps()
    org = o.combined.origin
    allp = np.concatenate([a, b]).astype(np.float64) - org - np.floor(np.array([3.5, 2.5, 0.5]) - org)
    idx = o.combined.index_map
    v = int(3 - org[0]) + int(2 - org[1]) * 16 + int(0 - org[2]) * 256
    m = o.combined.metrics[idx[v]]
    np.testing.assert_allclose(m[:3], allp.mean(axis=0), rtol=1e-5)
    cov = np.cov(allp.T, bias=True)
    np.testing.assert_allclose(m[[3, 4, 5, 6, 7, 8]], cov[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]], rtol=1e-4,
                               atol=1e-6)
    assert m[9] == 75


@pytest.mark.parametrize("k,carried", [(10, True), (11, False)])
def test_carry_over_needs_at_most_ten_passes(k, carried):
    """gvom.py:1058: a voxel occupied in the previous combined map stays occupied only
    while the current buffer saw it free with at most 10 ray passes."""
    params = (1.0, 1.0, 16, 8, 1, 0.5, 0.5, 0.5, 0.3, 2.0, 0.5, 1.0, 1, 1)
    o = OracleGvom(*params)
    ego = (8.5, 8.5, 4.5)                              # origin (0, 0, 0), ego voxel (8, 8, 4)
    o.Process_pointcloud(np.array([[11.5, 8.5, 4.5]], np.float32), ego)   # occupies (11, 8, 4)
    o.combine_maps()
    o.Process_pointcloud(np.full((k, 3), [13.5, 8.5, 4.5], np.float32), ego)  # k rays through (11, 8, 4)
    o.combine_maps()
    v = 11 + 8 * 16 + 4 * 256
    val = int(o.combined.index_map[v])
    if carried:
        assert val >= 0
    else:
        assert val == -1 - k


# -- 2D passes -------------------------------------------------------------------------------
def test_tilted_plane_slope_and_normalised_residual():
    """z = a x + b y: slope = atan(a), atan(b); the residual uses coefficients divided by
    m = sqrt(a^2 + b^2 + 1) (gvom.py:786-792), so an interior 3x3 fit of spacing s has
    mse = (1 - 1/m)^2 (a^2 + b^2) 2 s^2 / 3."""
    xyr, a, b = 0.4, 0.3, -0.2
    XY = 8
    x, y = np.meshgrid(np.arange(XY), np.arange(XY), indexing="ij")
    h = a * x * xyr + b * y * xyr + 1.0
    sx, sy, rough, delta = go.slope_guess(h, np.full((XY, XY), -1000.0), xyr)
    m = math.sqrt(a * a + b * b + 1)
    np.testing.assert_allclose(sx[1:-1, 1:-1], math.atan(a), rtol=1e-12)
    np.testing.assert_allclose(sy[1:-1, 1:-1], math.atan(b), rtol=1e-12)
    mse = (1 - 1 / m) ** 2 * (a * a + b * b) * 2 * xyr * xyr / 3
    np.testing.assert_allclose(rough[1:-1, 1:-1], math.log(mse), rtol=1e-9)
    assert np.all(delta == 0)


def test_flat_plane_roughness_zero_and_slope_defaults():
    XY = 6
    h = np.full((XY, XY), 2.5)
    sx, sy, rough, _ = go.slope_guess(h, np.full((XY, XY), -1000.0), 0.4)
    assert np.all(rough == 0.0) and np.all(sx == 0.0) and np.all(sy == 0.0)
    h2 = np.full((XY, XY), -1000.0)
    h2[0, 1:4] = 1.0                                   # collinear at x = 0: det == 0 exactly -> defaults
    h2[5, 5] = 1.0                                     # isolated: < 3 points -> defaults
    sx, sy, rough, _ = go.slope_guess(h2, np.full((XY, XY), -1000.0), 0.4)
    assert rough[0, 2] == -1.0 and sx[0, 2] == 0.0 and rough[5, 5] == -1.0


def _guess(height_cells, inferred0, XY=40, c=(20, 20)):
    h = np.full((XY, XY), -1000.0)
    for (dx, dy), v in height_cells.items():
        h[c[0] + dx, c[1] + dy] = v
    f = np.full((XY, XY), -1000.0)
    f[c] = inferred0
    return go.slope_guess(h, f, 0.4)[3][c]


def test_guess_height_four_neighbours():
    d = _guess({(1, 0): 1.0, (-1, 0): 2.0, (0, 1): 1.5, (0, -1): 0.5}, 0.8)
    assert d == 2.0 - 0.5


def test_guess_height_y_nh_is_gated_on_x_nh():
    # -x found, -y never found within 15: y_nh = -1000 enters min_h (gvom.py:704-706)
    d = _guess({(1, 0): 1.0, (-1, 0): 2.0, (0, 1): 1.5}, 0.8)
    assert d == 2.0 - (-1000.0)
    # -x never found, -y found: y_nh is ignored
    d = _guess({(1, 0): 1.0, (0, 1): 1.5, (0, -1): -3.0}, 0.8)
    assert d == 1.5 - 1.0


def test_guess_height_loop_ignores_x_p_done():
    # -x, +y, -y all found at ring 1, so the loop stops before +x reaches ring 3
    # (gvom.py:619 tests x_n_done twice and never x_p_done).
    d = _guess({(3, 0): 9.0, (-1, 0): 1.0, (0, 1): 1.2, (0, -1): 1.1}, 0.5)
    assert d == 1.2 - 1.0      # the +x cell at 9.0 would have made it 8.0


def test_guess_height_one_sided_scan_ranges():
    # +x scans dy in [-i, i): (1, +1) is outside ring 1's range and +x never sees it;
    # +y scans dx in (-i, i] and finds it; -x at (-1, +1) is inside (-i, i] and found.
    d = _guess({(1, 1): 3.0, (-1, 1): 0.5, (0, -1): 1.0}, 1.0)
    assert d == 3.0 - 0.5
    # -y scans dx in [-i, i): (-1, -1) is found at ring 1 by -y, +x finds (1, -1).
    d = _guess({(1, -1): 2.0, (-1, -1): 0.25, (-1, 0): 0.75}, 1.0)
    assert d == 2.0 - 0.25


def test_positive_obstacle_density_window():
    """__make_positive_obstacle_map (gvom.py:530-555): density over occupied voxels with
    hit > 10 in [floor((h+0.5)/zr - oz) + 1, floor((h+2)/zr - oz)], truncated to int."""
    params = (1.0, 0.5, 8, 16, 1, 0.1, 0.5, 0.5, 0.3, 2.0, 0.1, 1.0, 0, 0)
    o = OracleGvom(*params)
    o.Process_pointcloud(np.array([[100.0, 100.0, 100.0]], np.float32), (4.5, 4.5, 4.25))   # only sets up a slot
    sl = o.slots[0]
    V = o.voxel_count
    idx = np.full(V, -1, np.int32)
    col = (2, 3)
    zs = [1, 3, 4, 7, 8]           # ground at z=1; voxels above
    hits = [5, 12, 30, 50, 40]
    tots = [5, 20, 40, 60, 400]
    for i, z in enumerate(zs):
        idx[col[0] + col[1] * 8 + z * 64] = i
    sl.index_map = idx
    sl.hit = np.array(hits, np.int32)
    sl.total = np.array(tots, np.int32)
    sl.min_h = np.full(5, 0.25, np.float32)
    sl.metrics = np.zeros((5, 10))
    sl.metrics[:, 9] = 1.0
    o.combine_maps()
    oz = sl.origin[2]
    h = (0.25 + 1 + oz) * 0.5
    lo = math.floor((h + 0.5) / 0.5 - oz) + 1
    hi = math.floor((h + 2.0) / 0.5 - oz)
    sel = [i for i, z in enumerate(zs) if lo <= z <= hi and hits[i] > 10]
    dens = sum(hits[i] for i in sel) / sum(tots[i] for i in sel)
    assert o.combined.height[col] == h
    assert o.combined.pos[col] == int(dens * 100)


@pytest.mark.parametrize("zs,expect", [((4, 1, 3, 2), 100),    # delta = (4 - 1) z_res = 1.5 > 0.5
                                       ((3, 2, 3, 2), 0)])     # delta = 0.5 exactly: not > 0.5
def test_negative_obstacle_threshold(zs, expect):
    """__make_negative_obstacle_map (gvom.py:506-512) on a hand-built slot: an unobserved
    column with a free voxel (an inferred height) whose four neighbours are observed at voxel
    heights zs (+x, -x, +y, -y); the guessed height difference is max - min of the neighbours'
    heights (gvom.py:593-714, the inferred height lies below them), 100 where it exceeds the
    threshold."""
    params = (1.0, 0.5, 12, 16, 1, 0.1, 0.5, 0.5, 0.3, 2.0, 0.1, 1.0, 0, 0)
    o = OracleGvom(*params)
    o.Process_pointcloud(np.array([[100.0, 100.0, 100.0]], np.float32), (6.5, 6.5, 4.25))   # sets up a slot
    sl = o.slots[0]
    XYm = 12
    idx = np.full(o.voxel_count, -1, np.int32)
    c = (5, 5)
    idx[c[0] + c[1] * XYm + 2 * XYm * XYm] = -4          # free, 3 passes: inferred height at z = 2
    for i, ((dx, dy), z) in enumerate(zip(((1, 0), (-1, 0), (0, 1), (0, -1)), zs)):
        idx[(c[0] + dx) + (c[1] + dy) * XYm + z * XYm * XYm] = i
    sl.index_map = idx
    sl.hit = np.full(4, 3, np.int32)
    sl.total = np.full(4, 3, np.int32)
    sl.min_h = np.full(4, 0.25, np.float32)
    sl.metrics = np.zeros((4, 10))
    sl.metrics[:, 9] = 1.0
    out = o.combine_maps()
    oz = sl.origin[2]
    assert o.combined.height[c] == -1000.0 and o.combined.inferred[c] == (2 + oz) * 0.5
    assert o.combined.delta[c] == (max(zs) - min(zs)) * 0.5
    assert out[2][c] == expect
