"""Host-array calls replayed from captured launch sequences (runs on an MI355X).

From the second scan call in a row with the same point count, dtype and source (host arrays,
or one device buffer with overlap off), a handle replays a captured launch sequence instead of
launching each kernel.  Handle A takes that path; handle B has stage timing on, which keeps
every call on the eager path.  A third test covers overlap mode's pending metrics: slot exports
right after a scan call wait for them.  Both see the same calls and must give the same maps, bit for bit, after every
combine, along a sequence that changes each thing a replay depends on: the point count, the
point dtype, the transform's presence and dtype, the launch settings, the buffers (reserve),
several scans before one combine, and a rejected call in between.  The last maps are also
compared with the oracle.
"""
import numpy as np
import pytest

import gvom_synth as synth
from oracle.gvom_oracle import OracleGvom
from parity import assert_equal, compare_combined, compare_maps

pytestmark = pytest.mark.gpu


def _same(a, b, label):
    assert (a is None) == (b is None), label
    if a is None:
        return
    for name, x, y in zip(("origin", "pos", "neg", "rough", "vis"), a, b):
        assert np.array_equal(x, y), f"{label}: {name} differs between the replayed and the eager handle"
    print(f"{label}: ok", flush=True)


def test_host_replay_matches_eager():
    import gvom
    import gvom_native as nat
    wl = synth.workload(2, frames=2)
    base = wl.steps[0][0]
    other = wl.steps[1][0]
    A = gvom.Gvom(*wl.map_args)
    B = gvom.Gvom(*wl.map_args)
    B.set_timing(True)
    o = OracleGvom(*wl.map_args)

    def proc(pts, ego, T):
        A.Process_pointcloud(pts, ego, T)
        B.Process_pointcloud(pts, ego, T)
        o.Process_pointcloud(pts, ego, T)

    def comb(label):
        # the oracle combines at the same points: a combine carries the previous map over
        out_a = A.combine_maps()
        _same(out_a, B.combine_maps(), label)
        return out_a, o.combine_maps()

    def ego(k):
        return base.ego + np.array([0.6 * k, -0.35 * k, 0.15 * (k % 3)])

    # same count: the scan stage is captured on the second call
    for k in range(4):
        proc(base.points, ego(k), base.transform)
        comb(f"frame {k}")
    # another count (eager), captured again on its second call, then back
    half = np.ascontiguousarray(base.points[: base.points.shape[0] // 2])
    proc(half, ego(4), base.transform)
    comb("half count")
    proc(half, ego(5), base.transform)
    comb("half count, replayed")
    proc(other.points, other.ego, other.transform)
    comb("other scan")
    proc(base.points, ego(6), base.transform)
    proc(base.points, ego(7), base.transform)
    comb("full count again, two scans")
    # float64 points, no transform, a float64 transform
    proc(base.points.astype(np.float64), ego(8), base.transform)
    proc(base.points.astype(np.float64), ego(9), None)
    comb("float64 points")
    proc(base.points.astype(np.float64), ego(10), base.transform.astype(np.float64))
    comb("float64 transform")
    # three scans before one combine: the two pinned staging buffers are reused in turn
    for k in range(11, 14):
        proc(base.points, ego(k), base.transform)
    comb("three scans, one combine")
    # a rejected call changes nothing on either handle
    for g in (A, B):
        with pytest.raises(nat.GvomError):
            g.Process_pointcloud(base.points, np.array([np.nan, 0.0, 1.0]), base.transform)
    proc(base.points, ego(14), base.transform)
    comb("after a rejected call")
    # new launch settings and new buffers: captured again
    for g in (A, B):
        g.set_raycast_config(256, 1, False, False, False)
    proc(base.points, ego(15), base.transform)
    proc(base.points, ego(16), base.transform)
    comb("direct ray cast")
    for g in (A, B):
        g.reserve(base.points.shape[0] * 2)
    proc(base.points, ego(17), base.transform)
    comb("after reserve")
    proc(base.points, ego(18), base.transform)
    out_a, out_o = comb("after reserve, replayed")
    # the replayed handle against the oracle
    compare_combined(A.combined_state(), o.combined, "replayed handle combined")
    compare_maps(A.maps_state(), o.combined, "replayed handle")
    assert_equal("origin_m", out_a[0], out_o[0])
    for name, a, b in zip(("pos", "neg", "vis"), (out_a[1], out_a[2], out_a[4]), (out_o[1], out_o[2], out_o[4])):
        assert_equal(name, a, b)


def test_device_replay_matches_eager():
    import gvom
    import gvom_native as nat
    wl = synth.workload(2, frames=2)
    base = wl.steps[0][0]
    other = wl.steps[1][0]
    A = gvom.Gvom(*wl.map_args)
    B = gvom.Gvom(*wl.map_args)
    B.set_timing(True)
    o = OracleGvom(*wl.map_args)
    d_base = nat.DeviceArray(base.points)
    d_other = nat.DeviceArray(other.points)
    d_copy = nat.DeviceArray(base.points)

    def dproc(buf, n, pts, ego, T):
        for g in (A, B):
            g.process_pointcloud_device(buf, n, ego, T)
        o.Process_pointcloud(pts, ego, T)

    def hproc(pts, ego, T):
        for g in (A, B):
            g.Process_pointcloud(pts, ego, T)
        o.Process_pointcloud(pts, ego, T)

    def comb(label):
        out_a = A.combine_maps()
        _same(out_a, B.combine_maps(), label)
        return out_a, o.combine_maps()

    def ego(k):
        return base.ego + np.array([0.5 * k, 0.3 * k, -0.1 * (k % 2)])

    n = base.points.shape[0]
    # one device buffer: captured on the second call, then replayed
    for k in range(3):
        dproc(d_base, n, base.points, ego(k), base.transform)
        comb(f"device frame {k}")
    # another buffer with the same points, then back: a new source each time
    dproc(d_copy, n, base.points, ego(3), base.transform)
    dproc(d_copy, n, base.points, ego(4), base.transform)
    comb("second device buffer")
    dproc(d_base, n, base.points, ego(5), base.transform)
    comb("first buffer again")
    # fewer points from the same buffer (another count)
    dproc(d_base, n // 3, base.points[: n // 3], ego(6), base.transform)
    dproc(d_base, n // 3, base.points[: n // 3], ego(7), None)
    comb("a third of the points")
    # host arrays in between, then the device buffer again
    hproc(base.points, ego(8), base.transform)
    hproc(base.points, ego(9), base.transform)
    dproc(d_other, other.points.shape[0], other.points, other.ego, other.transform)
    dproc(d_other, other.points.shape[0], other.points, other.ego + 0.25, other.transform)
    comb("host arrays, then another scan's buffer")
    # overlap on: eager calls (the stage forks from the previous commit), same maps
    for g in (A, B):
        g.set_overlap(True)
    dproc(d_base, n, base.points, ego(10), base.transform)
    dproc(d_base, n, base.points, ego(11), base.transform)
    out_a, out_o = comb("overlap on")
    compare_combined(A.combined_state(), o.combined, "device replay combined")
    compare_maps(A.maps_state(), o.combined, "device replay")
    assert_equal("origin_m", out_a[0], out_o[0])
    for name, a, b in zip(("pos", "neg", "vis"), (out_a[1], out_a[2], out_a[4]), (out_o[1], out_o[2], out_o[4])):
        assert_equal(name, a, b)


def test_overlap_getters_wait_for_metrics():
    """With overlap on, a scan stage's neighbourhood metrics may still run after its call
    returns; a slot export right after it (no combine in between) must see them finished, as a
    serial handle does, and so must the maps of the combines that follow."""
    import gvom
    import gvom_native as nat
    wl = synth.workload(2, frames=2)
    scans = [st[0] for st in wl.steps]
    dev = [nat.DeviceArray(s.points) for s in scans]
    serial = gvom.Gvom(*wl.map_args)
    ovl = gvom.Gvom(*wl.map_args)
    ovl.set_overlap(True)
    for k in range(5):
        s = scans[k % 2]
        ego = s.ego + np.array([0.35 * k, -0.2 * k, 0.05 * k])
        for g in (serial, ovl):
            g.process_pointcloud_device(dev[k % 2], s.points.shape[0], ego, s.transform)
        for slot in range(serial.buffer_size):
            a, b = serial.slot_state(slot), ovl.slot_state(slot)
            assert (a is None) == (b is None), f"scan {k} slot {slot}"
            if a is not None:
                for name in a:
                    assert np.array_equal(a[name], b[name]), f"scan {k} slot {slot} {name}"
        if k % 2:
            _same(serial.combine_maps(), ovl.combine_maps(), f"combine after scan {k}")
    print("overlapped slot exports and combines equal the serial handle's", flush=True)
