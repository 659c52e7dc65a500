"""HIP path vs the oracle on the same seeded scans (runs on an MI355X).

Every scan slot, every combined map and every 2D map is compared after each call:
integers bit-exactly per voxel, floats within rtol 1e-4 / atol 1e-5.
"""
import numpy as np
import pytest

import gvom_synth as synth
from oracle.gvom_oracle import OracleGvom
from parity import compare_combined, compare_maps, compare_scan, assert_equal, assert_close

pytestmark = pytest.mark.gpu


def _gvom(args):
    import gvom
    return gvom.Gvom(*args)


def _run_and_compare(args, steps, label, check_every=True):
    g = _gvom(args)
    o = OracleGvom(*args)
    for k, step in enumerate(steps):
        for s in step:
            g.Process_pointcloud(s.points, s.ego, s.transform)
            o.Process_pointcloud(s.points, s.ego, s.transform)
            if check_every:
                slot = o.last_buffer_index
                compare_scan(g.slot_state(slot), o.slots[slot], f"{label} step {k} slot {slot}")
        out_g = g.combine_maps()
        out_o = o.combine_maps()
        compare_combined(g.combined_state(), o.combined, f"{label} step {k} combined")
        compare_maps(g.maps_state(), o.combined, f"{label} step {k}")
        assert_equal(f"{label} origin_m", out_g[0], out_o[0])
        for name, a, b in zip(("pos", "neg", "vis"), (out_g[1], out_g[2], out_g[4]), (out_o[1], out_o[2], out_o[4])):
            assert_equal(f"{label} step {k} {name}", a, b)
        assert_close(f"{label} step {k} roughness", out_g[3], out_o[3])
        print(f"{label} step {k}: cells {len(o.combined.hit)} ok", flush=True)
    return g, o


def test_cfg1_plain():
    wl = synth.workload(1)
    _run_and_compare(wl.map_args, wl.steps, "cfg1")


def test_cfg2_four_frames():
    wl = synth.workload(2)
    _run_and_compare(wl.map_args, wl.steps, "cfg2")


def test_cfg3_large_map():
    """512 x 512 x 64 voxels (16.8M per table), 236k-point 128 x 2048 scans, buffer 8, the
    sensor moving and turning along a curved path: all eight frames, so the 8-slot ring fills
    (gvom.py:198-216, 242-298); every new slot, combined map and 2D map checked."""
    wl = synth.workload(3, frames=8)
    _run_and_compare(wl.map_args, wl.steps, "cfg3")


def test_f64_points_and_transform():
    wl = synth.workload(1)
    s = wl.steps[0][0]
    T = np.eye(4)
    T[:3, :3] = np.array([[0.8, -0.6, 0.0], [0.6, 0.8, 0.0], [0.0, 0.0, 1.0]])
    T[:3, 3] = [0.3, -0.2, 0.05]
    pts = s.points.astype(np.float64) @ T[:3, :3]   # arbitrary sensor-frame cloud
    step = [[synth.ScanInput(pts, s.ego, T, 0)]]
    _run_and_compare(wl.map_args, step, "f64")


def test_no_transform_and_mixed_dtypes():
    wl = synth.workload(1)
    s = wl.steps[0][0]
    steps = [[synth.ScanInput(s.points, s.ego, None, 0)],
             [synth.ScanInput(s.points, s.ego + np.array([0.5, 0.3, 0.0]), s.transform.astype(np.float64), 0)]]
    args = list(wl.map_args)
    args[4] = 2
    _run_and_compare(tuple(args), steps, "mixed")


def test_combine_before_process_returns_none(capsys):
    wl = synth.workload(1)
    g = _gvom(wl.map_args)
    assert g.combine_maps() is None
    assert "ERROR: No data in buffer" in capsys.readouterr().out
    assert g.make_debug_voxel_map() is None


def test_empty_and_degenerate_scans():
    wl = synth.workload(1)
    s = wl.steps[0][0]
    empty = np.zeros((0, 3), np.float32)
    near = np.full((100, 3), 0.1, np.float32)           # all inside min_distance
    nanpts = s.points.copy()
    nanpts[::7] = np.nan
    far = np.full((50, 3), 1e6, np.float32)             # all outside the map
    args = list(wl.map_args)
    args[4] = 2
    steps = [[synth.ScanInput(empty, s.ego, s.transform, 0)],
             [synth.ScanInput(near, s.ego, s.transform, 0)],
             [synth.ScanInput(nanpts, s.ego, s.transform, 0)],
             [synth.ScanInput(far, s.ego, s.transform, 0)],
             [synth.ScanInput(s.points, s.ego + np.array([40.0, 0.0, 0.0]), s.transform, 0)]]
    _run_and_compare(tuple(args), steps, "edge")


@pytest.mark.parametrize("variant", [(256, 1, 0), (128, 1, 4)])
def test_odd_map_sizes(variant):
    """A map whose sides are not multiples of the kernels' tiles (37 x 37 columns, 13 z
    levels, buffer 3): partial 16 x 16 column tiles in the 2D pass, partial 4 x 4 ray-cast
    tiles, z columns shorter than a wavefront.  Three frames of the small scan, both ray
    casts."""
    import gvom
    wl = synth.workload(1)
    args = (0.8, 0.4, 37, 13, 3) + tuple(wl.map_args[5:])
    steps = [[synth.ScanInput(wl.steps[0][0].points, wl.steps[0][0].ego + np.array([0.9 * k, -0.4 * k, 0.0]),
                              wl.steps[0][0].transform, k)] for k in range(3)]
    g = gvom.Gvom(*args)
    g.set_raycast_config(*variant[:2], False, False, bool(variant[2] & 4))
    o = OracleGvom(*args)
    for k, step in enumerate(steps):
        for s in step:
            g.Process_pointcloud(s.points, s.ego, s.transform)
            o.Process_pointcloud(s.points, s.ego, s.transform)
            slot = o.last_buffer_index
            compare_scan(g.slot_state(slot), o.slots[slot], f"odd {variant} step {k} slot {slot}")
        out_g = g.combine_maps()
        out_o = o.combine_maps()
        compare_combined(g.combined_state(), o.combined, f"odd {variant} step {k} combined")
        compare_maps(g.maps_state(), o.combined, f"odd {variant} step {k}")
        for name, a, b in zip(("pos", "neg", "vis"), (out_g[1], out_g[2], out_g[4]), (out_o[1], out_o[2], out_o[4])):
            assert_equal(f"odd {variant} step {k} {name}", a, b)
        # OccupancyGrid payloads on a side that is not a multiple of 4 (byte stores)
        from oracle import ros_ref
        ms = g.maps_state()
        grids = g.occupancy_grids()
        exp = ros_ref.occupancy_payloads(ms["pos"], ms["neg"], ms["rough"], ms["vis"])
        for name, v in exp.items():
            assert_equal(f"odd {variant} step {k} {name}", grids[name], v)
        print(f"odd map {variant} step {k}: cells {len(o.combined.hit)} ok", flush=True)


@pytest.mark.parametrize("variant", [(128, 1, 4), (256, 1, 0)])
def test_vertical_and_lateral_motion(variant):
    """Ego heights and positions that move the map origin by a few voxels up, down and
    sideways between frames, so that every source of the combine is read shifted in x, y
    and z; 256 x 256 x 64 map, buffer 4, five frames, each checked against the oracle."""
    import gvom
    wl = synth.workload(2)
    s0 = wl.steps[0][0]
    dz = [0.0, 0.5, -0.7, 1.3, -2.1]
    steps = [[synth.ScanInput(s0.points, s0.ego + np.array([0.45 * k, -0.3 * k, dz[k]]), s0.transform, k)]
             for k in range(5)]
    g = gvom.Gvom(*wl.map_args)
    g.set_raycast_config(*variant[:2], False, False, bool(variant[2] & 4))
    o = OracleGvom(*wl.map_args)
    for k, step in enumerate(steps):
        for s in step:
            g.Process_pointcloud(s.points, s.ego, s.transform)
            o.Process_pointcloud(s.points, s.ego, s.transform)
            slot = o.last_buffer_index
            compare_scan(g.slot_state(slot), o.slots[slot], f"motion {variant} step {k} slot {slot}")
        out_g = g.combine_maps()
        out_o = o.combine_maps()
        compare_combined(g.combined_state(), o.combined, f"motion {variant} step {k} combined")
        compare_maps(g.maps_state(), o.combined, f"motion {variant} step {k}")
        assert_equal(f"motion {variant} step {k} origin_m", out_g[0], out_o[0])
        print(f"motion {variant} step {k}: cells {len(o.combined.hit)} ok", flush=True)


def test_jumps_beyond_the_map():
    """Ego jumps larger than the map's z extent (12.8 m) and its xy half-width (51.2 m), so
    that whole sources of the combine fall outside the new map window, then back; 256 x 256 x
    64 map, buffer 4, every slot, combined map and 2D map checked against the oracle."""
    import gvom
    wl = synth.workload(2)
    s0 = wl.steps[0][0]
    jumps = [(0.0, 0.0, 0.0), (0.0, 0.0, 14.0), (0.0, 0.0, -3.0), (60.0, -55.0, 0.0), (0.0, 0.0, 0.0),
             (0.3, -0.2, 20.5)]
    steps = [[synth.ScanInput(s0.points, s0.ego + np.array(d), s0.transform, k)] for k, d in enumerate(jumps)]
    g = gvom.Gvom(*wl.map_args)
    o = OracleGvom(*wl.map_args)
    for k, step in enumerate(steps):
        for s in step:
            g.Process_pointcloud(s.points, s.ego, s.transform)
            o.Process_pointcloud(s.points, s.ego, s.transform)
            slot = o.last_buffer_index
            compare_scan(g.slot_state(slot), o.slots[slot], f"jumps step {k} slot {slot}")
        out_g = g.combine_maps()
        out_o = o.combine_maps()
        compare_combined(g.combined_state(), o.combined, f"jumps step {k} combined")
        compare_maps(g.maps_state(), o.combined, f"jumps step {k}")
        assert_equal(f"jumps step {k} origin_m", out_g[0], out_o[0])
        print(f"jumps step {k}: cells {len(o.combined.hit)} ok", flush=True)
