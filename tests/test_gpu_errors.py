"""Host-side validation through the C-ABI on a live handle.  Every call below is rejected
before any kernel is launched, raises GvomError (with the library's message where it sets
one), and leaves the handle's ring, slots, combined map and 2D maps exactly as a handle that
never saw the bad calls has them."""
import numpy as np
import pytest

import gvom_synth as synth

pytestmark = pytest.mark.gpu


def _same(a, b, label):
    for k in a:
        assert np.array_equal(a[k], b[k]), (label, k)


def test_rejected_calls_leave_the_handle_intact():
    import gvom
    import gvom_native as nat
    from gvom_torch import MapLoss
    wl = synth.workload(2)   # 256 x 256 x 64, buffer 4
    scans = [st[0] for st in wl.steps]
    g = gvom.Gvom(*wl.map_args)
    ref = gvom.Gvom(*wl.map_args)
    for h in (g, ref):
        h.Process_pointcloud(scans[0].points, scans[0].ego, scans[0].transform)
        h.combine_maps()

    def expect(fn, match):
        with pytest.raises(nat.GvomError, match=match):
            fn()
        print(f"rejected: {match}", flush=True)

    s = scans[1]
    # launch configuration
    expect(lambda: g.set_raycast_config(100, 1), "gvom call failed")
    expect(lambda: g.set_raycast_config(128, 0), "gvom call failed")
    expect(lambda: g.set_raycast_config(128, 1, use_box=True, tiled=True), "gvom call failed")
    # batch larger than the ring
    pts5 = np.concatenate([s.points] * 5)
    off5 = np.arange(6) * s.points.shape[0]
    expect(lambda: g.process_batch(pts5, off5, np.stack([s.ego] * 5), np.stack([s.transform] * 5)),
           "number of scans must be in")
    # ego outside the domain
    expect(lambda: g.Process_pointcloud(s.points, np.array([1.0e7, 0.0, 0.0]), s.transform),
           "ego position must be finite")
    expect(lambda: g.Process_pointcloud(s.points, np.array([np.nan, 0.0, 0.0]), s.transform),
           "ego position must be finite")
    # state export of a slot that does not exist
    expect(lambda: g.slot_state(4), "slot out of range")
    # backward without a forward (on a loss handle)
    expect(lambda: MapLoss(0.4, 0.2, 256, 64).backward(1.0), "needs a preceding gvom_diff_forward")
    # capture rules
    g.set_timing(True)
    expect(lambda: g.graph_begin(4), "disable timing before graph capture")
    g.set_timing(False)
    dev = nat.DeviceArray(s.points)   # allocated before the capture (no allocation inside one)
    g.graph_begin(4)
    expect(lambda: g.graph_begin(4), "a capture is already in progress")
    expect(lambda: g.Process_pointcloud(s.points, s.ego, s.transform), "only device-pointer process calls")
    expect(lambda: g.combine_maps(), "capture gvom_combine_maps_device instead")
    expect(lambda: g.set_overlap(True), "cannot change overlap while capturing")
    g.process_pointcloud_device(dev, s.points.shape[0], s.ego, s.transform)   # captured, never replayed
    g.combine_maps_device()
    gr = g.graph_end()
    del gr
    expect(lambda: g.graph_end(), "no capture in progress")

    assert g.ring_state() == ref.ring_state()
    for k in range(1, 4):
        sk = scans[k]
        for h in (g, ref):
            h.Process_pointcloud(sk.points, sk.ego, sk.transform)
            h.combine_maps()
        assert g.ring_state() == ref.ring_state()
        for slot in range(4):
            a, b = g.slot_state(slot), ref.slot_state(slot)
            assert (a is None) == (b is None)
            if a is not None:
                _same(a, b, f"frame {k} slot {slot}")
        _same(g.combined_state(), ref.combined_state(), f"frame {k} combined")
        _same(g.maps_state(), ref.maps_state(), f"frame {k} maps")
    print("handle unchanged by the rejected calls", flush=True)


def test_graph_refused_after_buffers_grow():
    """A captured sequence refers to the handle's device buffers: after an eager call that
    reallocates them (a scan larger than any before) its replay is refused with an error,
    and a sequence captured afterwards replays normally."""
    import gvom
    import gvom_native as nat
    wl = synth.workload(2)
    s = wl.steps[0][0]
    g = gvom.Gvom(*wl.map_args)
    dev = nat.DeviceArray(s.points)
    g.process_pointcloud_device(dev, s.points.shape[0], s.ego, s.transform)
    g.combine_maps_device()
    g.sync()
    g.graph_begin(1)
    g.process_pointcloud_device(dev, s.points.shape[0], s.ego, s.transform)
    g.combine_maps_device()
    gr = g.graph_end()
    big = np.concatenate([s.points] * 4)   # more points than the work buffers hold
    g.Process_pointcloud(big, s.ego, s.transform)
    with pytest.raises(nat.GvomError, match="reallocated after this sequence was captured"):
        gr.launch(s.ego[None, :], s.transform[None])
    g.sync()
    g.graph_begin(1)
    g.process_pointcloud_device(dev, s.points.shape[0], s.ego, s.transform)
    g.combine_maps_device()
    gr2 = g.graph_end()
    gr2.launch(s.ego[None, :], s.transform[None])
    g.sync()
    print("stale graph refused, recaptured graph replayed", flush=True)
