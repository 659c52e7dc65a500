"""Second, independent restatement of scripts/gvom.py in plain Python -- TEST INFRASTRUCTURE.

Slow (Python loops over points and voxels); used only on tiny maps to pin the C
oracle (gvom_oracle.c), and by tests/golden/make_golden.py to write golden vectors.
It follows the reference kernels statement by statement, with every value typed as
Numba's CUDA target types it: numpy float32 scalars where the reference holds f32,
float64 elsewhere.  NumPy >= 2 keeps a float32 scalar float32 when combined with a
Python float (NEP 50) whereas Numba promotes to f64, so every Python-float operand
is wrapped in np.float64 explicitly.

State layout is the reference's: dense int32 index maps with v = x + y*xy + z*xy*xy,
compact per-cell arrays with ids in voxel order.
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32
F64 = np.float64


def pymax(a, b):   # numba lowering of max(a, b): b > a ? b : a
    return b if b > a else a


def pymin(a, b):
    return b if b < a else a


class PyRef:
    def __init__(self, xy_res, z_res, xy, z, buf, min_d, pos_thr, neg_thr, slope_thr, robot_h, robot_r, g2l,
                 ex, ez):
        self.xyr, self.zr = F64(xy_res), F64(z_res)
        self.XY, self.Z, self.B = int(xy), int(z), int(buf)
        self.min_d = F64(min_d)
        self.pos_thr, self.neg_thr, self.slope_thr = F64(pos_thr), F64(neg_thr), F64(slope_thr)
        self.robot_h, self.robot_r, self.g2l = F64(robot_h), F64(robot_r), F64(g2l)
        self.ex, self.ez = int(ex), int(ez)
        self.V = self.XY * self.XY * self.Z
        self.slots = [None] * self.B
        self.buffer_index = 0
        self.last_buffer_index = 0
        self.last = None
        self.ego = (F64(0), F64(0), F64(0))

    def lin(self, x, y, z):
        return int(x) + int(y) * self.XY + int(z) * self.XY * self.XY

    # -- __transform_pointcloud (gvom.py:1122-1138) ------------------------------------
    @staticmethod
    def transform(p, T):
        pt = []
        for r in range(3):
            acc = p[0] * T[r][0] + p[1] * T[r][1] + p[2] * T[r][2] + T[r][3]
            pt.append(acc)
        return [type(p[0])(v) for v in pt]

    def valid(self, p):
        if not all(math.isfinite(float(v)) for v in p):
            return False
        d2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2]
        return not (F64(d2) < self.min_d * self.min_d)

    # -- Process_pointcloud (gvom.py:105-220) --------------------------------------------
    def process(self, points, ego, T=None):
        pts = np.asarray(points)
        dt = F32 if pts.dtype == np.float32 else F64
        ego = tuple(F64(v) for v in ego)
        self.ego = ego
        origin = (F64(math.floor(float(ego[0] / self.xyr - F64(self.XY) / F64(2)))),
                  F64(math.floor(float(ego[1] / self.xyr - F64(self.XY) / F64(2)))),
                  F64(math.floor(float(ego[2] / self.zr - F64(self.Z) / F64(2)))))
        world = []
        for row in pts:
            p = [dt(v) for v in row]
            if T is not None:
                Tm = np.asarray(T)
                tt = F32 if (Tm.dtype == np.float32 and dt is F32) else F64
                p = [tt(v) for v in p]
                Tr = [[tt(Tm[r][c]) for c in range(4)] for r in range(3)]
                p = [dt(v) for v in self.transform(p, Tr)]
            world.append(p)
        hit = {}
        total = {}
        XY, Z = self.XY, self.Z
        for p in world:
            if not self.valid(p):
                continue
            ix = math.floor(float(F64(p[0]) / self.xyr - origin[0]))
            iy = math.floor(float(F64(p[1]) / self.xyr - origin[1]))
            iz = math.floor(float(F64(p[2]) / self.zr - origin[2]))
            if 0 <= ix < XY and 0 <= iy < XY and 0 <= iz < Z:
                v = self.lin(ix, iy, iz)
                hit[v] = hit.get(v, 0) + 1
                total[v] = total.get(v, 0) + 1
            for v in self.ray_voxels(p, ego, origin):
                total[v] = total.get(v, 0) + 1
        index_map = np.empty(self.V, np.int32)
        ids = {}
        for v in range(self.V):
            if hit.get(v, 0) > 0:
                ids[v] = len(ids)
                index_map[v] = ids[v]
            else:
                index_map[v] = -total.get(v, 0) - 1
        C = len(ids)
        hit_c = np.zeros(C, np.int32)
        tot_c = np.zeros(C, np.int32)
        for v, i in ids.items():
            hit_c[i] = hit[v]
            tot_c[i] = total[v]
        metrics = self.metrics(world, origin, index_map, C)
        min_h = np.full(C, F32(1.0), np.float32)
        for p in world:
            if not self.valid(p):
                continue
            qx = F64(p[0]) / self.xyr - origin[0]
            qy = F64(p[1]) / self.xyr - origin[1]
            qz = F64(p[2]) / self.zr - origin[2]
            ix, iy, iz = math.floor(float(qx)), math.floor(float(qy)), math.floor(float(qz))
            if not (0 <= ix < XY and 0 <= iy < XY and 0 <= iz < Z):
                continue
            lz = qz - F64(iz)
            i = index_map[self.lin(ix, iy, iz)]
            min_h[i] = pymin(min_h[i], F32(lz))
        slot = dict(index_map=index_map, hit=hit_c, total=tot_c, metrics=metrics, min_h=min_h,
                    origin=np.array(origin))
        self.slots[self.buffer_index] = slot
        self.last_buffer_index = self.buffer_index
        self.buffer_index = (self.buffer_index + 1) % self.B
        return slot

    def ray_voxels(self, p, ego, origin):
        """Voxels counted by one ray of __point_2_map (gvom.py:1174-1231), in step order."""
        out = []
        pt = [F32(ego[0] / self.xyr), F32(ego[1] / self.xyr), F32(ego[2] / self.zr)]
        end = [F32(F64(p[0]) / self.xyr), F32(F64(p[1]) / self.xyr), F32(F64(p[2]) / self.zr)]
        s = [end[k] - pt[k] for k in range(3)]
        rl = F32(np.sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]))
        if not (F64(0) < F64(rl) - F64(1)):
            return out
        s = [s[k] / rl for k in range(3)]
        smax = pymax(abs(s[0]), pymax(abs(s[1]), abs(s[2])))
        si = 0
        if smax == abs(s[1]):
            si = 1
        if smax == abs(s[2]):
            si = 2
        length = F64(0)
        direction = s[si] / abs(s[si])
        sizes = (self.XY, self.XY, self.Z)
        while length < F64(rl) - F64(1):
            pt[si] = F32(pt[si] + direction)
            pt[(si + 1) % 3] = F32(pt[(si + 1) % 3] + s[(si + 1) % 3] / abs(s[si]))
            pt[(si + 2) % 3] = F32(pt[(si + 2) % 3] + s[(si + 2) % 3] / abs(s[si]))
            idx = []
            for k in range(3):
                f = math.floor(float(F64(pt[k]) - origin[k]))
                if f < 0 or f >= sizes[k]:
                    return out
                idx.append(f)
            out.append(self.lin(*idx))
            length = length + abs(F64(1.0) / F64(s[si]))
        return out

    def metrics(self, world, origin, index_map, C):
        """__calculate_mean / __normalize_mean / __calculate_covariance /
        __normalize_covariance (gvom.py:1250-1387)."""
        m = np.zeros((C, 10), np.float64)
        XY, Z, ex, ez = self.XY, self.Z, self.ex, self.ez
        contrib = []   # (cell, local) pairs in point order, reused by the covariance pass
        for p in world:
            if not self.valid(p):
                continue
            q = (F64(p[0]) / self.xyr - origin[0], F64(p[1]) / self.xyr - origin[1],
                 F64(p[2]) / self.zr - origin[2])
            b = [math.floor(float(v)) for v in q]
            for x in range(b[0] - ex, b[0] + 1 + ex):
                if x < 0 or x >= XY:
                    continue
                for y in range(b[1] - ex, b[1] + 1 + ex):
                    if y < 0 or y >= XY:
                        continue
                    for z in range(b[2] - ez, b[2] + 1 + ez):
                        if z < 0 or z >= Z:
                            continue
                        i = index_map[self.lin(x, y, z)]
                        if i < 0:
                            continue
                        loc = (q[0] - F64(x), q[1] - F64(y), q[2] - F64(z))
                        contrib.append((i, loc))
                        m[i, 0] += loc[0]
                        m[i, 1] += loc[1]
                        m[i, 2] += loc[2]
                        m[i, 9] += 1.0
        for i in range(C):
            for j in range(3):
                m[i, j] = m[i, j] / m[i, 9]
        for i, loc in contrib:
            d = (loc[0] - m[i, 0], loc[1] - m[i, 1], loc[2] - m[i, 2])
            m[i, 3] += d[0] * d[0]
            m[i, 4] += d[0] * d[1]
            m[i, 5] += d[0] * d[2]
            m[i, 6] += d[1] * d[1]
            m[i, 7] += d[1] * d[2]
            m[i, 8] += d[2] * d[2]
        for i in range(C):
            for j in range(6):
                m[i, j + 3] = 0.0 if m[i, 9] <= 0 else m[i, j + 3] / m[i, 9]
        return m

    # -- combine_maps (gvom.py:222-393) ----------------------------------------------------
    def _shift_iter(self, co, oo):
        XY, Z = self.XY, self.Z
        d = [co[k] - oo[k] for k in range(3)]
        for z in range(Z):
            for y in range(XY):
                for x in range(XY):
                    u = (F64(x) + d[0], F64(y) + d[1], F64(z) + d[2])
                    if u[0] >= XY or u[1] >= XY or u[2] >= Z or u[0] < 0 or u[1] < 0 or u[2] < 0:
                        continue
                    yield self.lin(x, y, z), self.lin(int(u[0]), int(u[1]), int(u[2]))

    def combine(self):
        newest = self.slots[self.last_buffer_index]
        if newest is None:
            return None
        co = newest["origin"]
        comb = np.full(self.V, -1, np.int64)
        counter = 0
        srcs = [s for s in self.slots if s is not None]
        for s in srcs:
            for v, u in self._shift_iter(co, s["origin"]):
                o = s["index_map"][u]
                if o >= 0 and comb[v] <= -1:
                    comb[v] = counter
                    counter += 1
                elif o < -1 and comb[v] <= -1:
                    comb[v] += o + 1
        L = self.last
        if L is not None:
            for v, u in self._shift_iter(co, L["origin"]):
                o = L["index_map"][u]
                if o >= 0 and -11 <= comb[v] <= -1:
                    comb[v] = counter
                    counter += 1
                elif o < -1 and comb[v] <= -1:
                    comb[v] += o + 1
        C = counter
        cm = np.zeros((C, 10), np.float32)
        chit = np.zeros(C, np.int64)
        ctot = np.zeros(C, np.int64)
        cmh = np.full(C, F32(1.0), np.float32)
        for s in srcs + ([L] if L is not None else []):
            for v, u in self._shift_iter(co, s["origin"]):
                ci, oi = comb[v], s["index_map"][u]
                if ci < 0 or oi < 0:
                    continue
                self._merge(cm[ci], s["metrics"][oi])
                chit[ci] += s["hit"][oi]
                ctot[ci] += s["total"][oi]
                cmh[ci] = pymin(cmh[ci], s["min_h"][oi])
        ev = np.zeros((C, 3), np.float32)
        for i in range(C):
            ev[i] = self.eigen(cm[i])
        res = dict(index_map=comb.astype(np.int32), hit=chit.astype(np.int32), total=ctot.astype(np.int32),
                   metrics=cm, min_h=cmh, eigen=ev, origin=np.array(co))
        self.maps(res)
        self.last = res
        return res

    @staticmethod
    def _merge(c, o):
        """One __combine_metrics update (gvom.py:921-977); c is a float32 row."""
        def prod(a, b):   # f32*f32 stays f32, anything with f64 is f64
            return a * b
        n = c[9] + o[9]
        mx = (prod(c[0], c[9]) + prod(o[0], o[9])) / n
        my = (prod(c[1], c[9]) + prod(o[1], o[9])) / n
        mz = (prod(c[2], c[9]) + prod(o[2], o[9])) / n
        mean = (mx, my, mz)
        new = []
        for k, (a, b) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))):
            val = (c[9] * c[3 + k] + o[9] * o[3 + k] + c[9] * (c[a] - mean[a]) * (c[b] - mean[b]) +
                   o[9] * (o[a] - mean[a]) * (o[b] - mean[b])) / n
            new.append(val)
        for k in range(6):
            c[3 + k] = new[k]
        c[0], c[1], c[2] = mx, my, mz
        c[9] = c[9] + o[9]

    @staticmethod
    def eigen(m):
        """__calculate_eigenvalues (gvom.py:1424-1487)."""
        xx, xy, xz, yy, yz, zz = (m[3], m[4], m[5], m[6], m[7], m[8])
        p1 = xy * xy + xz * xz + yz * yz
        q = F64(xx + yy + zz) / F64(3.0)
        e = np.zeros(3, np.float32)
        if p1 == 0:
            e[0] = pymax(xx, pymax(yy, zz))
            e[2] = pymin(xx, pymin(yy, zz))
            e[1] = F64(3.0) * q - F64(e[0]) - F64(e[2])
            return e
        p2 = (F64(xx) - q) * (F64(xx) - q) + (F64(yy) - q) * (F64(yy) - q) + (F64(zz) - q) * (F64(zz) - q) + F64(2.0) * F64(p1)
        p = math.sqrt(p2 / 6.0)
        B = [(F64(xx) - q) / p, F64(xy) / p, F64(xz) / p, (F64(yy) - q) / p, F64(yz) / p, (F64(zz) - q) / p]
        r = B[0] * (B[3] * B[5] - B[4] * B[4]) - B[1] * (B[1] * B[5] - B[4] * B[2]) + B[2] * (B[1] * B[4] - B[3] * B[2])
        r = r / 2
        if r <= -1:
            phi = math.pi / 3.0
        elif r >= 1:
            phi = 0.0
        else:
            phi = math.acos(r) / 3.0
        e[0] = q + 2.0 * p * math.cos(phi)
        e[2] = q + 2.0 * p * math.cos(phi + (2.0 * math.pi / 3.0))
        e[1] = F64(3.0) * q - F64(e[0]) - F64(e[2])
        return e

    def maps(self, res):
        XY, Z = self.XY, self.Z
        co = res["origin"]
        idx = res["index_map"]
        height = np.full((XY, XY), -1000.0)
        inferred = np.full((XY, XY), -1000.0)
        ego = self.ego
        for x in range(XY):
            for y in range(XY):
                xp = ((co[0] + F64(x)) * self.xyr) - ego[0]
                yp = ((co[1] + F64(y)) * self.xyr) - ego[1]
                if xp * xp + yp * yp <= self.robot_r * self.robot_r:
                    height[x, y] = ego[2] - self.g2l
                for z in range(Z):
                    i = idx[self.lin(x, y, z)]
                    if i >= 0:
                        height[x, y] = ((F64(res["min_h"][i]) + F64(z)) + co[2]) * self.zr
                        break
                for z in range(Z):
                    if idx[self.lin(x, y, z)] < -1:
                        inferred[x, y] = (F64(z) + co[2]) * self.zr
                        break
        sx, sy, rough, delta = slope_guess(height, inferred, self.xyr)
        pos = np.zeros((XY, XY), np.int32)
        neg = np.zeros((XY, XY), np.int32)
        vis = (height > -1000).astype(np.int32)
        for x in range(XY):
            for y in range(XY):
                if delta[x, y] > self.neg_thr:
                    neg[x, y] = 100
                if math.sqrt(sx[x, y] * sx[x, y] + sy[x, y] * sy[x, y]) >= self.slope_thr:
                    pos[x, y] = 100
                    continue
                lo = int(math.floor((height[x, y] + self.pos_thr) / self.zr - co[2])) + 1
                hi = int(math.floor((height[x, y] + self.robot_h) / self.zr - co[2]))
                if not (0 <= lo < Z) or not (0 <= hi < Z):
                    continue
                dens, n = 0.0, 0.0
                for z in range(lo, hi + 1):
                    i = idx[self.lin(x, y, z)]
                    if i >= 0 and res["hit"][i] > 10:
                        n += float(res["total"][i])
                        dens += float(res["hit"][i])
                if n > 0.0:
                    dens /= n
                pos[x, y] = int(dens * 100)
        res.update(height=height, inferred=inferred, slope_x=sx, slope_y=sy, rough=rough, delta=delta, pos=pos,
                   neg=neg, vis=vis)


def slope_guess(height, inferred, xyr):
    """__calculate_slope (gvom.py:718-805) and __guess_height (593-714) on 2D maps."""
    XY = height.shape[0]
    sx = np.zeros((XY, XY))
    sy = np.zeros((XY, XY))
    rough = np.full((XY, XY), -1.0)
    delta = np.zeros((XY, XY))
    for x0 in range(XY):
        for y0 in range(XY):
            pts = []
            for x in range(max(0, x0 - 1), min(XY, x0 + 2)):
                for y in range(max(0, y0 - 1), min(XY, y0 + 2)):
                    if height[x, y] > -1000:
                        pts.append((F64(x) * xyr, F64(y) * xyr, F64(height[x, y])))
            if len(pts) < 3:
                continue
            mx = my = mz = F64(0.0)
            for px, py, pz in pts:
                mx += px
                my += py
                mz += pz
            mx /= F64(len(pts))
            my /= F64(len(pts))
            mz /= F64(len(pts))
            xx = xy = xz = yy = yz = F64(0.0)
            for px, py, pz in pts:
                xx += (px - mx) * (px - mx)
                xy += (px - mx) * (py - my)
                xz += (px - mx) * (pz - mz)
                yy += (py - my) * (py - my)
                yz += (py - my) * (pz - mz)
            det = xx * yy - xy * xy
            if det == 0.0:
                continue
            a0 = (yy * xz - xy * yz) / det
            a1 = (xx * yz - xy * xz) / det
            m = math.sqrt(a0 * a0 + a1 * a1 + 1)
            a0 /= m
            a1 /= m
            err = F64(0.0)
            for px, py, pz in pts:
                e = (pz - mz) - (a0 * (px - mx) + a1 * (py - my))
                err += e * e
            err /= F64(len(pts))
            if err > 0:
                err = math.log(err)
            rough[x0, y0] = err
            sx[x0, y0] = math.atan2(a0, 1.0 / m)
            sy[x0, y0] = math.atan2(a1, 1.0 / m)
    for x0 in range(XY):
        for y0 in range(XY):
            if height[x0, y0] > -1000 or inferred[x0, y0] == -1000.0:
                continue
            done = {"xp": False, "xn": False, "yp": False, "yn": False}
            hv = {"xp": -1000.0, "xn": -1000.0, "yp": -1000.0, "yn": -1000.0}
            i = 0
            while i < 15 and not (done["xn"] and done["xn"] and done["yp"] and done["yn"]):
                i += 1
                scans = (("xp", x0 + i < XY, [(x0 + i, y0 + d) for d in range(-i, i)]),
                         ("xn", x0 - i >= 0, [(x0 - i, y0 + d) for d in range(-i + 1, i + 1)]),
                         ("yp", y0 + i < XY, [(x0 + d, y0 + i) for d in range(-i + 1, i + 1)]),
                         ("yn", y0 - i >= 0, [(x0 + d, y0 - i) for d in range(-i, i)]))
                for key, inside, cells in scans:
                    if done[key]:
                        continue
                    if not inside:
                        done[key] = True
                        continue
                    for cx, cy in cells:
                        if not (0 <= cx < XY and 0 <= cy < XY):
                            continue
                        if height[cx, cy] > -1000:
                            hv[key] = height[cx, cy]
                            done[key] = True
                            break
            mn, mxh = 1000.0, inferred[x0, y0]
            if hv["xp"] > -1000:
                mn, mxh = pymin(hv["xp"], mn), pymax(hv["xp"], mxh)
            if hv["xn"] > -1000:
                mn, mxh = pymin(hv["xn"], mn), pymax(hv["xn"], mxh)
            if hv["yp"] > -1000:
                mn, mxh = pymin(hv["yp"], mn), pymax(hv["yp"], mxh)
            if hv["xn"] > -1000:   # gvom.py:704 gates y_nh on x_nh
                mn, mxh = pymin(hv["yn"], mn), pymax(hv["yn"], mxh)
            dh = mxh - mn
            if dh > 0:
                delta[x0, y0] = dh
    return sx, sy, rough, delta
