"""Parity helpers: compare the HIP path's exported state with the oracle's.

Integer fields are compared bit-exactly per voxel (after mapping cell ids back to
voxel coordinates); float fields within rtol 1e-4 or atol 1e-5 (metres or voxel
units), the tolerance the path is specified to.
"""
import numpy as np

from oracle.gvom_oracle import canonical_scan

RTOL = 1e-4
ATOL = 1e-5


def close(a, b, rtol=RTOL, atol=ATOL):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    ok = np.isclose(a, b, rtol=rtol, atol=atol) | both_nan
    return ok


def assert_close(name, a, b, rtol=RTOL, atol=ATOL):
    ok = close(a, b, rtol, atol)
    if not ok.all():
        bad = np.argwhere(~ok)[:5]
        a = np.asarray(a)
        b = np.asarray(b)
        detail = [(tuple(i), a[tuple(i)], b[tuple(i)]) for i in bad]
        raise AssertionError(f"{name}: {(~ok).sum()} of {ok.size} differ beyond tolerance, e.g. {detail}")


def assert_equal(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape or not np.array_equal(a, b):
        if a.shape != b.shape:
            raise AssertionError(f"{name}: shape {a.shape} vs {b.shape}")
        bad = np.argwhere(a != b)[:5]
        detail = [(tuple(i), a[tuple(i)], b[tuple(i)]) for i in bad]
        raise AssertionError(f"{name}: {(a != b).sum()} of {a.size} differ, e.g. {detail}")


def compare_scan(gpu, orc, label=""):
    """gpu: dict from Gvom.slot_state; orc: oracle ScanSlot."""
    g = canonical_scan(gpu["index_map"], gpu["hit"], gpu["total"], gpu["min_h"], gpu["metrics"])
    o = canonical_scan(orc.index_map, orc.hit, orc.total, orc.min_h, orc.metrics)
    assert_equal(f"{label} origin", gpu["origin"], orc.origin)
    assert_equal(f"{label} cells", g["cells"], o["cells"])
    assert_equal(f"{label} occupancy", g["occ"], o["occ"])
    assert_equal(f"{label} passes", g["passes"], o["passes"])
    assert_equal(f"{label} hit", g["hit"], o["hit"])
    assert_equal(f"{label} total", g["total"], o["total"])
    occ = o["occ"]
    assert_equal(f"{label} min_h", g["min_h"][occ], o["min_h"][occ])
    assert_close(f"{label} metrics", g["metrics"][occ], o["metrics"][occ])


def compare_combined(gpu, orc, label=""):
    g = canonical_scan(gpu["index_map"], gpu["hit"], gpu["total"], gpu["min_h"], gpu["metrics"])
    o = canonical_scan(orc.index_map, orc.hit, orc.total, orc.min_h, orc.metrics)
    assert_equal(f"{label} origin", gpu["origin"], orc.origin)
    assert_equal(f"{label} cells", g["cells"], o["cells"])
    assert_equal(f"{label} occupancy", g["occ"], o["occ"])
    assert_equal(f"{label} passes", g["passes"], o["passes"])
    assert_equal(f"{label} hit", g["hit"], o["hit"])
    assert_equal(f"{label} total", g["total"], o["total"])
    occ = o["occ"]
    assert_equal(f"{label} min_h", g["min_h"][occ], o["min_h"][occ])
    assert_close(f"{label} metrics", g["metrics"][occ], o["metrics"][occ])
    ge = canonical_scan(gpu["index_map"], gpu["hit"], gpu["total"], None, gpu["eigen"])["metrics"][occ]
    oe = canonical_scan(orc.index_map, orc.hit, orc.total, None, orc.eigen)["metrics"][occ]
    # eigenvalues: 1e-4 relative, or 1e-5 of the trace for the smallest (f32-stored, near-planar voxels)
    tr = np.abs(oe).sum(axis=1, keepdims=True)
    ok = close(ge, oe) | (np.abs(ge - oe) <= 1e-5 * np.maximum(tr, 1e-12))
    if not ok.all():
        bad = np.argwhere(~ok)[:5]
        raise AssertionError(f"{label} eigen: {(~ok).sum()} differ, e.g. {[(tuple(i), ge[tuple(i)], oe[tuple(i)]) for i in bad]}")


def compare_maps(gpu_maps, orc, label=""):
    for k, ref in (("height", orc.height), ("inferred", orc.inferred), ("delta", orc.delta)):
        assert_close(f"{label} {k}", gpu_maps[k], ref)
    for k, ref in (("slope_x", orc.slope_x), ("slope_y", orc.slope_y), ("rough", orc.rough)):
        assert_close(f"{label} {k}", gpu_maps[k], ref)
    for k, ref in (("pos", orc.pos), ("neg", orc.neg), ("vis", orc.vis)):
        assert_equal(f"{label} {k}", gpu_maps[k], ref)
