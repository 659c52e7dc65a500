"""Pins of the FP64 differentiable-loss restatement (oracle/diff_ref.py) by what the
mathematics fixes, on the CPU:

- closed forms: boxes of 8 corner points whose covariance and smallest eigenvalue are known,
  and a two-column height variance;
- the analytic gradient of lambda_min for such a box (2/n (v . d_i) v, scaled by 1/res);
- rigid-motion invariances of the loss with the voxel assignment frozen: a translation, and
  a rotation about the vertical axis (the xy resolution is isotropic), leave it unchanged,
  so its gradient has zero net force and zero vertical torque; a tilt does change it (z is
  scaled by a different resolution), so the tests are not vacuous.
"""
import math

import numpy as np
import pytest
import torch

from oracle import diff_ref

XYR, ZR, XY, Z = 0.4, 0.2, 8, 8
RES = np.array([XYR, XYR, ZR])


def _box(voxel, half, yaw=0.0):
    """World points at the 8 corners of a box centred in `voxel`, half sizes `half` in
    voxel units, rotated by `yaw` about z (isotropic in xy, so also a world rotation)."""
    c, s = math.cos(yaw), math.sin(yaw)
    pts = []
    for sx in (-1, 1):
        for sy in (-1, 1):
            for sz in (-1, 1):
                d = np.array([sx * half[0], sy * half[1], sz * half[2]])
                d = np.array([c * d[0] - s * d[1], s * d[0] + c * d[1], d[2]])
                pts.append((np.asarray(voxel, float) + 0.5 + d) * RES)
    return np.array(pts)


def _assign(world):
    valid = np.ones(len(world), bool)
    return diff_ref.assign(world, valid, np.zeros(3), XYR, ZR, XY, Z)


def test_closed_form_two_boxes():
    a = _box((1, 1, 1), (0.3, 0.2, 0.1))
    b = _box((5, 5, 3), (0.15, 0.35, 0.25), yaw=math.radians(30))
    world = np.concatenate([a, b])
    asg = _assign(world)
    L, Lr, Lh = (float(x) for x in diff_ref.loss(torch.from_numpy(world), asg))
    assert Lr == pytest.approx(0.1 ** 2 + 0.15 ** 2, rel=1e-12)
    # lowest points: (1 + 0.5 - 0.1) * 0.2 and (3 + 0.5 - 0.25) * 0.2; population variance
    h = np.array([1.4 * ZR, 3.25 * ZR])
    assert Lh == pytest.approx(0.1 * h.var(), rel=1e-12)
    assert L == pytest.approx(Lr + Lh, rel=1e-15)


def test_box_gradient_closed_form():
    half = (0.3, 0.2, 0.1)
    world = _box((2, 3, 4), half)
    asg = _assign(world)
    w = torch.from_numpy(world).requires_grad_(True)
    _, Lr, _ = diff_ref.loss(w, asg)
    Lr.backward()
    g = w.grad.numpy()
    d = world / RES - (np.array([2, 3, 4]) + 0.5)
    # lambda_min = v^T C v with v = e_z: dL/dq_i = 2/n (v . d_i) v, q = world / res
    expect = np.zeros_like(world)
    expect[:, 2] = 2.0 / 8 * d[:, 2] / ZR
    np.testing.assert_allclose(g, expect, rtol=1e-10, atol=1e-14)


@pytest.fixture(scope="module")
def cloud():
    rng = np.random.default_rng(7)
    # a tilted, bumpy patch over the map plus a few blobs: many voxels with mixed neighbourhoods
    xy = rng.uniform(0.3, XY * XYR - 0.3, size=(1500, 2))
    z = 0.6 + 0.15 * xy[:, 0] + 0.1 * np.sin(2.0 * xy[:, 1]) + rng.normal(0, 0.03, 1500)
    blobs = rng.normal([1.5, 1.5, 1.0], [0.2, 0.2, 0.2], size=(200, 3))
    world = np.concatenate([np.column_stack([xy, z]), blobs])
    return world, _assign(world)


def _rot(axis, ang):
    c, s = math.cos(ang), math.sin(ang)
    if axis == 2:
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    return np.array([[1.0, 0, 0], [0, c, -s], [0, s, c]])


def _L(world, asg):
    with torch.no_grad():
        return float(diff_ref.loss(torch.from_numpy(world), asg)[0])


def test_loss_rigid_invariances(cloud):
    world, asg = cloud
    L0 = _L(world, asg)
    assert _L(world + np.array([0.37, -0.21, 0.11]), asg) == pytest.approx(L0, rel=1e-10)
    ctr = world.mean(0)
    yawed = (world - ctr) @ _rot(2, 0.3).T + ctr
    assert _L(yawed, asg) == pytest.approx(L0, rel=1e-10)
    tilted = (world - ctr) @ _rot(0, 0.3).T + ctr
    assert abs(_L(tilted, asg) - L0) > 1e-3 * abs(L0)


def test_gradient_net_force_and_vertical_torque_vanish(cloud):
    world, asg = cloud
    w = torch.from_numpy(world).requires_grad_(True)
    diff_ref.loss(w, asg)[0].backward()
    g = w.grad.numpy()
    scale = np.abs(g).sum()
    force = g.sum(0)
    torque = np.cross(world, g).sum(0)
    lever = np.abs(np.cross(world, g)).sum()
    assert np.abs(force).max() <= 1e-12 * scale
    assert abs(torque[2]) <= 1e-12 * lever
    assert np.abs(torque[:2]).max() > 1e-3 * lever
