"""The ROS node's conversions (CPU): the oracle's restatements pinned by hand-worked
cases, and the product's host-side transform and PointCloud2 layout extraction.

Payload values worked by hand from gvom_ros.py:142-162 with density_threshold 50 and
roughness range [-10, 0]: roughness r maps to ((clip(r) - 10) / 10) * 100, i.e. -100
at r = 0, -150 at r = -5, -200 at r <= -10; numpy's float64 -> int8 cast on x86-64
truncates to int32 and keeps the low byte: -150 -> 106, -200 -> 56, -110 -> -110.
"""
import math

import numpy as np
import pytest

import gvom_native as nat
from oracle import ros_ref
from ros_msgs import FLOAT32, FLOAT64, INT16, UINT8, pack_cloud


def _maps():
    # [x, y] arrays, x = row index
    pos = np.array([[0, 50], [51, 100]], np.int32)
    neg = np.array([[100, 0], [0, 0]], np.int32)
    rough = np.array([[-5.0, 0.0], [-1.0, -12.0]])
    vis = np.array([[1, 0], [1, 1]], np.int32)
    return pos, neg, rough, vis


def test_payloads_hand_worked():
    pos, neg, rough, vis = _maps()
    p = ros_ref.occupancy_payloads(pos, neg, rough, vis)
    # Fortran order: [x0y0, x1y0, x0y1, x1y1]
    assert p["hard_obstacle_map"].tolist() == [100, 100, 0, 100]
    assert p["soft_obstacle_map"].tolist() == [0, 0, 100, 0]
    assert p["ground_certainty_map"].tolist() == [100, 100, 0, 100]
    assert p["negative_obstacle_map"].tolist() == [100, 0, 0, 0]
    assert p["roughness_map"].tolist() == [106, -110, -100, 56]
    assert all(v.dtype == np.int8 for v in p.values())


def test_int8_cast_rule():
    """The cast the payloads end with, as the kernels restate it: truncate to int32
    (out of range or NaN -> INT32_MIN), keep the low byte."""
    v = np.array([-150.0, -200.0, 300.0, 127.9, -128.5, -0.5, 2.0 ** 31 + 5, -2.0 ** 31 - 5, 1e300, np.nan,
                  np.inf, -np.inf, -2147483648.7, 2147483647.2])

    def rule(x):
        t = -2 ** 31 if not (-2147483649.0 < x < 2147483648.0) else int(math.trunc(x))
        b = t & 0xFF
        return b - 256 if b > 127 else b

    with np.errstate(invalid="ignore"):
        got = v.astype(np.int8).tolist()
    assert got == [rule(x) for x in v]


def test_transform_closed_forms():
    from gvom_node import transform_from_translation_rotation as tfr
    t = [1.5, -2.0, 0.25]
    assert np.array_equal(tfr(t, [0, 0, 0, 1]), np.array([[1, 0, 0, 1.5], [0, 1, 0, -2.0], [0, 0, 1, 0.25],
                                                          [0, 0, 0, 1]], float))
    s = math.sqrt(0.5)
    Rz = tfr([0, 0, 0], [0, 0, s, s])[:3, :3]
    np.testing.assert_allclose(Rz, [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-15)
    Rx = tfr([0, 0, 0], [1, 0, 0, 0])[:3, :3]      # 180 deg about x
    np.testing.assert_allclose(Rx, [[1, 0, 0], [0, -1, 0], [0, 0, -1]], atol=1e-15)
    # scale invariance, and the tiny-quaternion identity rule
    q = np.array([0.1, -0.3, 0.2, 0.9])
    np.testing.assert_allclose(tfr(t, 3 * q), tfr(t, q), atol=1e-15)
    assert np.array_equal(tfr(t, [1e-9, 0, 0, 1e-9])[:3, :3], np.eye(3))


def test_transform_matches_restatement_bitwise():
    from gvom_node import transform_from_translation_rotation as tfr
    rng = np.random.default_rng(7)
    for _ in range(200):
        q = rng.normal(size=4) * rng.uniform(0.1, 3)
        t = rng.normal(size=3) * 10
        a, b = tfr(t, q), ros_ref.from_translation_rotation(t, q)
        assert np.array_equal(a, b)
        R = a[:3, :3]
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)


def _grid_points(n, seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-20, 20, size=(n, 3))


@pytest.mark.parametrize("layout,big", [
    ({"x": (0, FLOAT32), "y": (4, FLOAT32), "z": (8, FLOAT32)}, False),
    ({"x": (0, FLOAT64), "y": (8, FLOAT64), "z": (16, FLOAT64)}, True),
    ({"z": (0, FLOAT32), "x": (4, FLOAT32), "y": (12, FLOAT32)}, False),
    ({"x": (0, INT16), "y": (2, INT16), "z": (4, UINT8)}, True),
])
def test_pointcloud2_restatement(layout, big):
    h, w = 3, 5
    pts = _grid_points(h * w)
    if layout["z"][1] == UINT8:
        pts[:, 2] = np.abs(pts[:, 2])
    if layout["x"][1] in (FLOAT32, FLOAT64):
        pts[[1, 7], 0] = np.nan
        pts[11, 2] = np.inf
    msg = pack_cloud(pts, h, w, layout, big=big)
    fields = {f.name: (f.offset, f.datatype) for f in msg.fields}
    got = ros_ref.pointcloud2_to_xyz(msg.data, h, w, msg.point_step, fields, big)
    typ = {FLOAT32: np.float32, FLOAT64: np.float64, INT16: np.int16, UINT8: np.uint8}
    exp = np.stack([pts[:, k].astype(typ[layout[n][1]]).astype(np.float64) for k, n in enumerate("xyz")], 1)
    exp = exp[np.isfinite(exp).all(1)]
    assert got.dtype == np.float64
    assert np.array_equal(got, exp)
    if layout["x"][1] in (FLOAT32, FLOAT64):
        assert got.shape[0] == h * w - 3


def test_cloud_layout_extraction():
    from gvom import cloud_layout
    msg = pack_cloud(_grid_points(8), 2, 4, {"x": (0, FLOAT32), "y": (4, FLOAT32), "z": (8, FLOAT32)},
                     row_pad=12)
    lay = cloud_layout(msg)
    assert isinstance(lay, nat.CloudLayout)
    assert (lay.height, lay.width, lay.point_step, lay.row_step) == (2, 4, msg.point_step, 4 * msg.point_step + 12)
    assert list(lay.offset) == [0, 4, 8] and list(lay.datatype) == [FLOAT32] * 3 and lay.is_bigendian == 0
    msg.fields = [f for f in msg.fields if f.name != "z"]
    with pytest.raises(ValueError):
        cloud_layout(msg)
