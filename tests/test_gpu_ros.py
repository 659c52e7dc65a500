"""The ROS node's input and output conversions on the GPU.

- occupancy_grids: the int8 payloads made on the device are bit-identical to the
  node's numpy expressions (oracle/ros_ref.py) applied to the same combined maps, for
  the default parameters and for parameters that reach the cast's edge cases.
- process_pointcloud2: PointCloud2 bytes gathered on the device leave a slot that is
  bit-identical to Process_pointcloud on ros_numpy's float64 xyz array (NaN rows
  removed), for float32 / float64 / big-endian / padded-row layouts, and matches the
  oracle.
- GvomNode: odometry + PointCloud2 + tf translation/quaternion through one timer tick
  gives the payloads the oracle path (ros_numpy + tf restatements + C oracle + node
  formatting) gives.
"""
import ctypes

import numpy as np
import pytest

import gvom_synth as synth
from oracle import ros_ref
from oracle.gvom_oracle import OracleGvom
from parity import assert_equal, compare_combined, compare_scan
from ros_msgs import FLOAT32, FLOAT64, odometry, pack_cloud

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg2():
    return synth.workload(2, frames=3)


def _run_frames(g, wl, frames):
    out = None
    for step in wl.steps[:frames]:
        for s in step:
            g.Process_pointcloud(s.points, s.ego, s.transform)
        out = g.combine_maps()
    return out


@pytest.mark.parametrize("params", [(50, -10, 0), (49.5, -3.0, 2.0), (0, -1e12, 1e12), (100, -5, -5),
                                    (-1, 0.5, -0.5)])
def test_occupancy_grids_match_node_expressions(cfg2, params):
    import gvom
    import gvom_native as nat
    g = gvom.Gvom(*cfg2.map_args)
    assert g.occupancy_grids(*params) is None
    last = _run_frames(g, cfg2, 3)
    ms = g.maps_state()
    got = g.occupancy_grids(*params)
    exp = ros_ref.occupancy_payloads(ms["pos"], ms["neg"], ms["rough"], ms["vis"], *params)
    for name in nat.GRID_NAMES:
        assert_equal(f"{name} {params}", got[name], exp[name])
    np.testing.assert_array_equal(got["origin"], last[0])
    # the payloads are those of combine_maps' own outputs
    hp = ros_ref.occupancy_payloads(last[1], last[2], last[3], last[4], *params)
    for name in nat.GRID_NAMES:
        assert_equal(f"{name} {params} from combine_maps outputs", got[name], hp[name])
    # device output (no combine in between): same bytes
    X = cfg2.map_args[2] ** 2
    dev = nat.DeviceArray(np.zeros(5 * X, np.int8))
    g.occupancy_grids(*params, device_out=dev)
    g.sync()
    back = np.empty(5 * X, np.int8)
    nat._hip_runtime().hipMemcpy(back.ctypes.data_as(ctypes.c_void_p), dev.ptr, back.nbytes, 2)
    assert_equal(f"device payloads {params}", back, np.concatenate([got[n] for n in nat.GRID_NAMES]))
    print(f"payloads {params}: ok, hard>0 cells {int((got['hard_obstacle_map'] != 0).sum())}", flush=True)


LAYOUTS = {
    "f32": ({"x": (0, FLOAT32), "y": (4, FLOAT32), "z": (8, FLOAT32)}, False, 0),
    "f64_big": ({"x": (0, FLOAT64), "y": (8, FLOAT64), "z": (16, FLOAT64)}, True, 0),
    "f32_shuffled_padded_rows": ({"z": (0, FLOAT32), "y": (4, FLOAT32), "x": (12, FLOAT32)}, False, 20),
}


def _message(scan, layout, big, row_pad, width=1024, seed=0):
    pts = scan.points.astype(np.float64)
    n = pts.shape[0] - pts.shape[0] % width
    pts = pts[:n].copy()
    rng = np.random.default_rng(seed)
    bad = rng.choice(n, size=n // 50, replace=False)
    pts[bad, rng.integers(0, 3, size=bad.size)] = np.nan
    pts[bad[:5], 0] = np.inf
    return pack_cloud(pts, n // width, width, layout, big=big, row_pad=row_pad), pts


@pytest.mark.parametrize("name", list(LAYOUTS))
def test_pointcloud2_equals_xyz_path(cfg2, name):
    import gvom
    import gvom_native as nat
    layout, big, pad = LAYOUTS[name]
    a = gvom.Gvom(*cfg2.map_args)
    b = gvom.Gvom(*cfg2.map_args)
    o = OracleGvom(*cfg2.map_args)
    for k, step in enumerate(cfg2.steps):
        s = step[0]
        msg, pts = _message(s, layout, big, pad, seed=k)
        if pad == 0:
            fields = {f.name: (f.offset, f.datatype) for f in msg.fields}
            xyz = ros_ref.pointcloud2_to_xyz(msg.data, msg.height, msg.width, msg.point_step, fields, big)
        else:   # ros_numpy reads rows back to back; the unpadded message is the reference input
            xyz = ros_ref.pointcloud2_to_xyz(*_unpadded(pts, msg, layout, big))
        T = s.transform.astype(np.float64)
        if k == 1:
            # device-resident message bytes
            raw = nat.DeviceArray(np.frombuffer(msg.data, np.uint8).copy())
            a.process_pointcloud2(msg, s.ego, T, device_data=raw)
            a.sync()
        else:
            a.process_pointcloud2(msg, s.ego, T)
        b.Process_pointcloud(xyz, s.ego, T)
        o.Process_pointcloud(xyz, s.ego, T)
        slot = b.ring_state()["last_buffer_index"]
        sa, sb = a.slot_state(slot), b.slot_state(slot)
        for key in sb:
            assert_equal(f"{name} frame {k} {key}", sa[key], sb[key])
        compare_scan(sa, o.slots[slot], f"{name} frame {k} vs oracle")
        print(f"{name} frame {k}: {xyz.shape[0]} of {msg.height * msg.width} points kept, slot identical", flush=True)
    a.combine_maps()
    o.combine_maps()
    compare_combined(a.combined_state(), o.combined, f"{name} combined vs oracle")


def _unpadded(pts, msg, layout, big):
    m = pack_cloud(pts, msg.height, msg.width, layout, big=big, row_pad=0)
    fields = {f.name: (f.offset, f.datatype) for f in m.fields}
    return m.data, m.height, m.width, m.point_step, fields, big


def test_node_timer_matches_oracle_path(cfg2):
    from gvom_node import GvomNode
    args = cfg2.map_args
    node = GvomNode(xy_resolution=args[0], z_resolution=args[1], width=args[2], height=args[3], buffer_size=args[4])
    layout, big, _ = LAYOUTS["f32"]
    s0 = cfg2.steps[0][0]
    msg, _ = _message(s0, layout, big, 0)
    assert node.cb_lidar(msg, [0, 0, 0], [0, 0, 0, 1]) is False      # no odometry yet
    assert node.cb_timer() is None
    o = OracleGvom(*args)
    rng = np.random.default_rng(11)
    for k, step in enumerate(cfg2.steps):
        s = step[0]
        msg, _ = _message(s, layout, big, 0, seed=100 + k)
        t = rng.normal(size=3)
        q = rng.normal(size=4)
        node.cb_odom(odometry(*s.ego))
        assert node.cb_lidar(msg, t, q)
        fields = {f.name: (f.offset, f.datatype) for f in msg.fields}
        xyz = ros_ref.pointcloud2_to_xyz(msg.data, msg.height, msg.width, msg.point_step, fields, big)
        o.Process_pointcloud(xyz, tuple(float(v) for v in s.ego), ros_ref.from_translation_rotation(t, q))
        out = node.cb_timer(debug=(k == len(cfg2.steps) - 1))
        origin, pos, neg, rough, vis = o.combine_maps()
        exp = ros_ref.occupancy_payloads(pos, neg, rough, vis)
        for topic, payload in out.grids.items():
            ref = exp["ground_certainty_map" if topic == "all_ground_certainty_map" else topic]
            assert_equal(f"step {k} {topic}", payload, ref)
        assert out.info.origin_position == (origin[0], origin[1], 0.0)
        assert out.info.width == out.info.height == args[2] and out.info.resolution == args[0]
    assert set(out.debug) == {"voxel", "height_map", "inferred_height_map"}
    hm = out.debug["height_map"]
    assert hm.dtype.names == ("x", "y", "z", "roughness", "slope_x", "slope_y", "slope", "obstacles")
    np.testing.assert_array_equal(hm["obstacles"], pos.flatten("F"))
    print("node timer payloads identical to the oracle path", flush=True)


def test_pointcloud2_layout_errors(cfg2):
    """Malformed layouts are refused with GVOM_EINVAL before anything reaches the device,
    and leave the ring untouched."""
    import gvom
    import gvom_native as nat
    g = gvom.Gvom(*cfg2.map_args)
    s = cfg2.steps[0][0]
    msg, _ = _message(s, *LAYOUTS["f32"])
    before = g.ring_state()

    def bad(**change):
        m = pack_cloud(np.zeros((8, 3)), 2, 4, {"x": (0, FLOAT32), "y": (4, FLOAT32), "z": (8, FLOAT32)})
        for k, v in change.items():
            setattr(m, k, v)
        return m

    cases = [bad(point_step=0), bad(row_step=4), bad(height=-1)]
    m = bad()
    m.fields[2].datatype = 9
    cases.append(m)
    m = bad()
    m.fields[0].offset = m.point_step - 2
    cases.append(m)
    for m in cases:
        with pytest.raises(nat.GvomError):
            g.process_pointcloud2(m, s.ego)
    assert g.ring_state() == before
    g.process_pointcloud2(msg, s.ego, s.transform.astype(np.float64))
    assert g.ring_state() != before

