"""FP64 reference of the differentiable map loss -- TEST INFRASTRUCTURE ONLY.

A float64 torch (CPU) re-expression of the loss the HIP kernels differentiate:

    L = sum over occupied voxels u of lambda_min(C_u) + w_h * Var_c(h_c)

with C_u the population covariance of {p/res - origin - u} over valid points whose
own voxel is within (+-ex, +-ex, +-ez) of u (gvom.py:1250-1387), h_c the world z of the
lowest point of the lowest occupied voxel of column c.  `Assignment` freezes which
points belong to which voxel and which points are the column argmins, so the loss is a
smooth function of the point coordinates: torch.autograd gives its exact FP64
gradient, and central finite differences of the same function check that gradient.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class Assignment:
    pair_u: np.ndarray       # (M,) occupied-voxel index (0..U-1) of each (voxel, point) pair
    pair_i: np.ndarray       # (M,) point index
    u_xyz: np.ndarray        # (U, 3) voxel coordinates
    col_pts: np.ndarray      # (K,) argmin point of each column with an occupied voxel
    origin: np.ndarray
    res: np.ndarray          # (3,) xy, xy, z resolution


def assign(world, valid, origin, xy_res, z_res, XY, Z, ex=1, ez=1) -> Assignment:
    res = np.array([xy_res, xy_res, z_res])
    q = world / res - origin
    f = np.floor(q).astype(np.int64)
    ok = valid.copy()
    inb = ok & (f[:, 0] >= 0) & (f[:, 0] < XY) & (f[:, 1] >= 0) & (f[:, 1] < XY) & (f[:, 2] >= 0) & (f[:, 2] < Z)
    lin = f[:, 0] + f[:, 1] * XY + f[:, 2] * XY * XY
    occ_lin = np.unique(lin[inb])
    vid = -np.ones(XY * XY * Z, np.int64)
    vid[occ_lin] = np.arange(len(occ_lin))
    u_xyz = np.stack([occ_lin % XY, (occ_lin // XY) % XY, occ_lin // (XY * XY)], 1)
    pu, pi = [], []
    idx = np.nonzero(ok)[0]
    for dx in range(-ex, ex + 1):
        for dy in range(-ex, ex + 1):
            for dz in range(-ez, ez + 1):
                n = f[idx] + np.array([dx, dy, dz])
                m = (n[:, 0] >= 0) & (n[:, 0] < XY) & (n[:, 1] >= 0) & (n[:, 1] < XY) & (n[:, 2] >= 0) & (n[:, 2] < Z)
                nl = n[m, 0] + n[m, 1] * XY + n[m, 2] * XY * XY
                v = vid[nl]
                keep = v >= 0
                pu.append(v[keep])
                pi.append(idx[m][keep])
    pair_u = np.concatenate(pu)
    pair_i = np.concatenate(pi)
    # lowest occupied voxel per column, its lowest point (first index on ties)
    col = occ_lin % (XY * XY)
    zs = occ_lin // (XY * XY)
    order = np.lexsort((zs, col))
    first = np.ones(len(order), bool)
    first[1:] = col[order][1:] != col[order][:-1]
    low_lin = set(occ_lin[order][first].tolist())
    pts_in = np.nonzero(inb)[0]
    best = {}
    for i in pts_in:
        L = int(lin[i])
        if L in low_lin:
            zc = world[i, 2]
            if L not in best or zc < best[L][0]:
                best[L] = (zc, i)
    col_pts = np.array([best[L][1] for L in sorted(best)], np.int64)
    return Assignment(pair_u, pair_i, u_xyz, col_pts, np.asarray(origin, np.float64), res)


def loss(world: torch.Tensor, a: Assignment, w_h=0.1, covariances=False):
    """world: (N, 3) float64 tensor (requires_grad for gradients).  Returns (L, sum of
    lambda_min, height term), plus the (U, 3, 3) covariances when asked."""
    res = torch.tensor(a.res, dtype=torch.float64)
    org = torch.tensor(a.origin, dtype=torch.float64)
    q = world / res - org
    pu = torch.from_numpy(a.pair_u)
    pi = torch.from_numpy(a.pair_i)
    U = a.u_xyz.shape[0]
    loc = q[pi] - torch.from_numpy(a.u_xyz).double()[pu]
    n = torch.zeros(U, dtype=torch.float64).index_add_(0, pu, torch.ones(len(pu), dtype=torch.float64))
    mean = torch.zeros(U, 3, dtype=torch.float64).index_add_(0, pu, loc) / n[:, None]
    d = loc - mean[pu]
    outer = d[:, :, None] * d[:, None, :]
    C = torch.zeros(U, 3, 3, dtype=torch.float64).index_add_(0, pu, outer) / n[:, None, None]
    lam = torch.linalg.eigvalsh(C)[:, 0]
    h = world[torch.from_numpy(a.col_pts), 2]
    var = ((h - h.mean()) ** 2).mean() if len(h) else torch.zeros((), dtype=torch.float64)
    if covariances:
        return lam.sum() + w_h * var, lam.sum(), w_h * var, C
    return lam.sum() + w_h * var, lam.sum(), w_h * var
