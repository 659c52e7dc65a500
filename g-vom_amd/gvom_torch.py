"""Autograd boundary of the differentiable scan->map loss (torch is used only here).

`MapLoss` owns a dedicated library handle (ring of one slot) and exposes the
hand-written HIP forward/backward of gvom_diff_forward / gvom_diff_backward:

    L = sum over occupied voxels of lambda_min(neighbourhood covariance)
      + w_height * variance over columns of the lowest point's height,

with voxel assignment, counts and argmin points fixed by the forward pass.
`map_loss(points, xi, T0, ego, loss)` is a torch.autograd.Function over float64 CUDA
points (sensor frame) and a 6-vector xi = (rho, phi) that perturbs the transform as
T = exp(xi^) T0; its xi-gradient is the left-perturbation gradient at T, exact at xi = 0
(refinement loops re-anchor T0 every step).
"""
from __future__ import annotations

import ctypes

import numpy as np

import gvom
import gvom_native as nat


def hat(w):
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def se3_exp(xi):
    """SE(3) exponential of xi = (rho, phi) -> 4x4 (float64)."""
    xi = np.asarray(xi, np.float64)
    rho, phi = xi[:3], xi[3:]
    th = float(np.linalg.norm(phi))
    K = hat(phi)
    if th < 1e-12:
        R = np.eye(3) + K
        V = np.eye(3) + 0.5 * K
    else:
        A = np.sin(th) / th
        B = (1 - np.cos(th)) / th ** 2
        Cc = (th - np.sin(th)) / th ** 3
        R = np.eye(3) + A * K + B * K @ K
        V = np.eye(3) + B * K + Cc * K @ K
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = V @ rho
    return T


class MapLoss:
    """Differentiable map loss on single scans, backed by the HIP kernels."""

    def __init__(self, xy_resolution, z_resolution, xy_size, z_size, min_distance=1.0, xy_eigen_dist=1,
                 z_eigen_dist=1, w_height=0.1, geometry_only=True):
        self.g = gvom.Gvom(xy_resolution, z_resolution, xy_size, z_size, 1, min_distance, 0.5, 0.5, 0.3, 2.0, 4.0,
                           1.0, xy_eigen_dist, z_eigen_dist)
        # the handle only evaluates the loss: no ray cast (pass counts are not part of it)
        nat.check(self.g._lib.gvom_set_diff_geometry_only(self.g._h, 1 if geometry_only else 0), self.g._h)
        self.w_height = float(w_height)
        self.n = 0

    def forward(self, points_ptr, n, T, ego):
        """points_ptr: device pointer to n x 3 float64 sensor-frame points."""
        T = np.ascontiguousarray(np.asarray(T, np.float64).reshape(4, 4))
        ego = np.ascontiguousarray(np.asarray(ego, np.float64).reshape(3))
        out = np.zeros(3)
        nat.check(self.g._lib.gvom_diff_forward(self.g._h, ctypes.c_void_p(int(points_ptr)), int(n), nat.ptr(T),
                                                nat.ptr(ego), self.w_height, nat.ptr(out)), self.g._h)
        self.n = int(n)
        return out

    def forward_async(self, points_ptr, n, T, ego, out_ptr, stream):
        """Stream-ordered forward: ordered after the work on `stream` (a HIP stream handle,
        0 = null stream), which is ordered after it; writes {L, L_rough, w Var} to the device
        array at out_ptr (3 float64).  No host synchronisation."""
        T = np.ascontiguousarray(np.asarray(T, np.float64).reshape(4, 4))
        ego = np.ascontiguousarray(np.asarray(ego, np.float64).reshape(3))
        nat.check(self.g._lib.gvom_diff_forward_async(self.g._h, ctypes.c_void_p(int(points_ptr)), int(n), nat.ptr(T),
                                                      nat.ptr(ego), self.w_height, ctypes.c_void_p(int(out_ptr)),
                                                      ctypes.c_void_p(int(stream))), self.g._h)
        self.n = int(n)

    def backward_async(self, dL_ptr, grad_points_ptr, grad_world_ptr, grad_xi_ptr, stream):
        """Stream-ordered backward; dL_ptr: device float64 scalar, grad_xi_ptr: device 6 float64."""
        def vp(x):
            return None if x is None else ctypes.c_void_p(int(x))
        nat.check(self.g._lib.gvom_diff_backward_async(self.g._h, vp(dL_ptr), vp(grad_points_ptr), vp(grad_world_ptr),
                                                       vp(grad_xi_ptr), ctypes.c_void_p(int(stream))), self.g._h)

    def backward(self, dL=1.0, grad_points_ptr=None, grad_world_ptr=None):
        gxi = np.zeros(6)
        gp = None if grad_points_ptr is None else ctypes.c_void_p(int(grad_points_ptr))
        gw = None if grad_world_ptr is None else ctypes.c_void_p(int(grad_world_ptr))
        nat.check(self.g._lib.gvom_diff_backward(self.g._h, float(dL), gp, gw, nat.ptr(gxi)), self.g._h)
        return gxi


def _function():
    import torch
    runtimes = nat.hip_runtime_files()
    if len(runtimes) > 1:
        # torch's CUDA tensors would come from one HIP runtime and the kernels run on another
        raise nat.GvomError("torch and libgvom_hip.so use different HIP runtimes in this process "
                            f"({', '.join(runtimes)}); import torch before gvom so that both bind to one")

    class _MapLossFn(torch.autograd.Function):
        """The library's stream-ordered entry points on torch's current stream: the loss is a
        0-d float64 tensor on the points' device, and neither direction synchronises the host
        (only xi's host copy, for the transform, does when xi lives on the device)."""

        @staticmethod
        def forward(ctx, points, xi, T0, ego, loss):
            if points.dtype != torch.float64 or not points.is_cuda or not points.is_contiguous():
                raise ValueError("points must be a contiguous float64 CUDA tensor (N, 3)")
            T = se3_exp(xi.detach().cpu().numpy()) @ np.asarray(T0, np.float64)
            stream = torch.cuda.current_stream(points.device).cuda_stream
            out = torch.empty(3, dtype=torch.float64, device=points.device)
            loss.forward_async(points.data_ptr(), points.shape[0], T, ego, out.data_ptr(), stream)
            ctx.loss = loss
            ctx.xi_device = xi.device
            ctx.xi_dtype = xi.dtype
            ctx.save_for_backward(points)
            return out[0].clone()

        @staticmethod
        def backward(ctx, gL):
            (points,) = ctx.saved_tensors
            gp = torch.empty_like(points) if ctx.needs_input_grad[0] else None
            gL = gL.to(device=points.device, dtype=torch.float64).contiguous()
            gxi = torch.empty(6, dtype=torch.float64, device=points.device)
            stream = torch.cuda.current_stream(points.device).cuda_stream
            ctx.loss.backward_async(gL.data_ptr(), None if gp is None else gp.data_ptr(), None, gxi.data_ptr(),
                                    stream)
            return gp, gxi.to(device=ctx.xi_device, dtype=ctx.xi_dtype), None, None, None

    return _MapLossFn


_FN = None


def map_loss(points, xi, T0, ego, loss: MapLoss):
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(points, xi, T0, ego, loss)


def extrinsic_evaluator(scan_points, poses, loss, ego=None, E_ref=None):
    """evaluate(E) -> (L, dL/dxi) for a sensor extrinsic E (sensor -> vehicle) shared by scans
    taken at known vehicle poses P_i (vehicle -> world): maps every scan with W_i = P_i E p,
    runs the HIP forward/backward on the union (identity transform) for dL/dW, and chains it to
    a left perturbation E <- exp(xi^) E: dL/drho = sum R_i^T g, dL/dphi = sum q x (R_i^T g)
    with q = E p (vehicle frame).

    With E_ref, the common part of the translation change, c = mean_i R_i (t_E - t_ref), is
    taken off every point: the loss is invariant to a common translation of all points but
    for where the points fall in the voxel grid, and that grid aliasing would otherwise pull
    the extrinsic's translation toward a lower-loss alignment with the grid.  The translation
    is then found from how the frames align with one another."""
    P = [np.asarray(p, np.float64) for p in poses]
    if ego is None:
        ego = np.mean([p[:3, 3] for p in P], axis=0)
    M = np.mean([p[:3, :3] for p in P], axis=0)
    t_ref = None if E_ref is None else np.asarray(E_ref, np.float64)[:3, 3].copy()
    n = int(sum(s.shape[0] for s in scan_points))
    dev = nat.DeviceArray(np.zeros((n, 3)))
    gdev = nat.DeviceArray(np.zeros((n, 3)))
    hip = nat._hip_runtime()
    eye = np.eye(4)

    def evaluate(E, need_grad=True):
        qs = [s @ E[:3, :3].T + E[:3, 3] for s in scan_points]                  # vehicle frame
        c = np.zeros(3) if t_ref is None else M @ (E[:3, 3] - t_ref)
        W = np.ascontiguousarray(np.concatenate([q @ p[:3, :3].T + (p[:3, 3] - c) for q, p in zip(qs, P)]))
        hip.hipMemcpy(dev.ptr, W.ctypes.data_as(ctypes.c_void_p), W.nbytes, 1)
        L = float(loss.forward(dev.ptr.value, n, eye, ego)[0])
        if not need_grad:
            return L, None
        loss.backward(1.0, None, gdev.ptr.value)
        gw = np.empty((n, 3))
        hip.hipMemcpy(gw.ctypes.data_as(ctypes.c_void_p), gdev.ptr, gw.nbytes, 2)
        grad = np.zeros(6)
        o = 0
        for q, p in zip(qs, P):
            gl = gw[o:o + q.shape[0]] @ p[:3, :3]                                # R_i^T g per point
            grad[:3] += gl.sum(0)
            grad[3:] += np.cross(q, gl).sum(0)
            o += q.shape[0]
        if t_ref is not None:
            # c moves with xi as dc = M (drho + dphi x t_E): every point's -c adds -(dc/dxi)^T G
            G = M.T @ gw.sum(0)
            grad[:3] -= G
            grad[3:] -= np.cross(E[:3, 3], G)
        return L, grad
    return evaluate


def lm_refine(evaluate, E0, steps=20, free=(1, 1, 0, 1, 1, 1), h=(5e-3, 5e-3, 5e-3, 1e-3, 1e-3, 1e-3),
              scale=(0.02, 0.005), radius0=1.0, radius_max=4.0):
    """Trust-region Levenberg-Marquardt refinement of a 4x4 rigid transform E under left
    perturbations xi = (rho, phi): E <- exp(xi^) E.

    evaluate(E, need_grad=True) -> (L, dL/dxi).  Each step builds the Hessian H of the free
    coordinates from central differences of the analytic gradient (steps h) and takes the
    minimiser of the quadratic model g.d + d.H.d / 2 within a trust region |d / s| <= radius
    (s = `scale`: metres for rho, radians for phi), found on the Levenberg-Marquardt path
    (H + mu diag(1/s^2)) d = -g.  The step is accepted when the loss falls; the radius halves
    when the loss fell by less than a quarter of the model's prediction and doubles (up to
    radius_max) when the model predicted well at the boundary.  Held coordinates (free = 0)
    never move: a loss invariant along a direction gives no information there.  Returns
    (E, [loss after each step, starting with the initial loss], [accepted steps])."""
    E = np.asarray(E0, np.float64).copy()
    fr = np.nonzero(np.asarray(free))[0]
    h = np.asarray(h, np.float64)
    sc = np.array([scale[0]] * 3 + [scale[1]] * 3, np.float64)[fr]
    radius = radius0
    L, g = evaluate(E)
    losses, moves = [L], []
    for _ in range(steps):
        H = np.zeros((6, 6))
        for j in fr:
            e = np.zeros(6)
            e[j] = h[j]
            gp = evaluate(se3_exp(e) @ E)[1]
            gm = evaluate(se3_exp(-e) @ E)[1]
            H[:, j] = (gp - gm) / (2 * h[j])
        Hf = H[np.ix_(fr, fr)]
        Hf = 0.5 * (Hf + Hf.T)
        # scaled coordinates u = d / s
        Hs = Hf * np.outer(sc, sc)
        gs = g[fr] * sc
        w, V = np.linalg.eigh(Hs)
        b = V.T @ gs

        def step(mu):
            return -(V @ (b / (w + mu)))
        lo = max(0.0, -w.min()) + 1e-12 * max(1.0, np.abs(w).max())
        u = step(lo)
        if np.linalg.norm(u) > radius:
            hi = lo + 1.0
            while np.linalg.norm(step(hi)) > radius:
                hi = lo + 2.0 * (hi - lo)
            for _b in range(60):
                mid = 0.5 * (lo + hi)
                if np.linalg.norm(step(mid)) > radius:
                    lo = mid
                else:
                    hi = mid
            u = step(hi)
        d = u * sc
        pred = -(g[fr] @ d + 0.5 * d @ Hf @ d)
        xi = np.zeros(6)
        xi[fr] = d
        En = se3_exp(xi) @ E
        Ln, gn = evaluate(En)
        rho = (L - Ln) / pred if pred > 0 else -1.0
        if Ln < L:
            E, L, g = En, Ln, gn
            moves.append(xi)
        else:
            moves.append(np.zeros(6))
        if rho < 0.25:
            radius *= 0.5
        elif rho > 0.75 and np.linalg.norm(u) > 0.99 * radius:
            radius = min(2.0 * radius, radius_max)
        losses.append(L)
    return E, losses, moves


def refine_extrinsic(scan_points, poses, E0, loss, steps=20, free=(1, 1, 0, 1, 1, 1), ego=None, **kw):
    """Refine a sensor extrinsic E (sensor -> vehicle) from several scans taken at known
    vehicle poses by minimising the map loss of their union (lm_refine over
    extrinsic_evaluator).  The loss is invariant to a common translation of all points; with
    poses that differ only by yaw, the vertical translation of E moves every frame alike and
    is held by default (free[2] = 0).  Returns (E, [loss per step])."""
    E, losses, _ = lm_refine(extrinsic_evaluator(scan_points, poses, loss, ego, E_ref=E0), E0, steps=steps,
                             free=free, **kw)
    return E, losses
