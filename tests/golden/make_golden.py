"""Writes tests/golden/*.npz: tiny seeded scan->map cases with the outputs of the
pure-Python restatement (oracle/pyref.py).  Inputs and outputs only; rerun with
`python tests/golden/make_golden.py` to regenerate.

Each case: map parameters, a list of frames (points, ego, transform), and after each
frame the scan slot and the combined map in the reference's format.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.pyref import PyRef  # noqa: E402


def scene(rng, ego, n_ground, n_box, box_c, dtype):
    g = rng.uniform(-3.0, 3.0, (n_ground, 2)) + ego[:2]
    gz = 0.05 * np.sin(g[:, 0] * 2.0) + rng.normal(0, 0.01, n_ground)
    ground = np.column_stack([g, gz])
    b = rng.uniform(-0.15, 0.15, (n_box, 3)) + np.asarray(box_c)
    b[:, 2] = rng.uniform(0.0, 1.2, n_box)
    far = rng.uniform(-8, 8, (12, 3)) + np.array([ego[0], ego[1], 0.0])   # some beyond the map
    pts = np.concatenate([ground, b, far])
    return pts.astype(dtype)


def rot(yaw, tx, ty, tz, dtype):
    c, s = np.cos(yaw), np.sin(yaw)
    T = np.eye(4)
    T[:3, :3] = [[c, -s, 0], [s, c, 0], [0, 0, 1]]
    T[:3, 3] = [tx, ty, tz]
    return T.astype(dtype)


CASES = {
    # name: (params, dtype, eigen dists, frames)
    "tiny_f32": dict(params=(0.5, 0.25, 12, 8, 2, 1.0, 0.5, 0.5, 0.3, 2.0, 1.5, 1.0, 1, 1), dtype=np.float32,
                     frames=3, seed=7),
    "tiny_f64": dict(params=(0.4, 0.2, 10, 8, 1, 0.8, 0.4, 0.4, 0.25, 1.5, 1.0, 0.8, 0, 1), dtype=np.float64,
                     frames=2, seed=8),
    "tiny_halo": dict(params=(0.5, 0.3, 8, 6, 3, 0.5, 0.5, 0.5, 0.3, 2.0, 1.0, 1.0, 2, 1), dtype=np.float32,
                      frames=4, seed=9),
}


def make_case(name, spec):
    rng = np.random.default_rng(spec["seed"])
    ref = PyRef(*spec["params"])
    out = {"params": np.array(spec["params"], np.float64)}
    for f in range(spec["frames"]):
        ego = np.array([0.6 * f, 0.1 * f, 1.0 + 0.05 * f])
        world = scene(rng, ego, 160, 40, ego + np.array([1.2, -0.8, 0.0]), np.float64)
        if f == 1:
            world[::17] = np.nan
        T = rot(0.2 * f + 0.1, ego[0], ego[1], 0.0, spec["dtype"])
        Ti = np.linalg.inv(T.astype(np.float64))
        sensor = (world @ Ti[:3, :3].T + Ti[:3, 3]).astype(spec["dtype"])
        use_T = None if (name == "tiny_halo" and f == 2) else T
        ref.process(sensor, ego, use_T)
        slot = ref.slots[ref.last_buffer_index]
        res = ref.combine()
        out[f"f{f}_points"] = sensor
        out[f"f{f}_ego"] = ego
        out[f"f{f}_T"] = T if use_T is not None else np.zeros((0,), spec["dtype"])
        for k in ("index_map", "hit", "total", "metrics", "min_h", "origin"):
            out[f"f{f}_slot_{k}"] = slot[k]
        for k in ("index_map", "hit", "total", "metrics", "min_h", "eigen", "origin", "height", "inferred",
                  "slope_x", "slope_y", "rough", "delta", "pos", "neg", "vis"):
            out[f"f{f}_comb_{k}"] = res[k]
    out["frames"] = np.array(spec["frames"])
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    return out


if __name__ == "__main__":
    for name, spec in CASES.items():
        o = make_case(name, spec)
        print(name, "cells per frame:", [int((o[f"f{f}_comb_index_map"] >= 0).sum()) for f in range(spec["frames"])])
