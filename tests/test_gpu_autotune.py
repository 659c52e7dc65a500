"""gvom_autotune leaves the handle's maps alone (runs on an MI355X).

The autotuner times ray-cast variants on one device-resident scan through the handle's own
work buffers and a private table, saving and restoring the device ring state around it.  After
it, the ring, every slot, the combined map and the 2D maps must be exactly what they were, and
the next frames (with the variant it picked) must equal those of a handle that never ran it.
"""
import numpy as np
import pytest

import gvom_synth as synth

pytestmark = pytest.mark.gpu


def _snapshot(g, B):
    snap = {"ring": g.ring_state(), "combined": g.combined_state(), "maps": g.maps_state()}
    for k in range(B):
        snap[f"slot{k}"] = g.slot_state(k)
    return snap


def _equal(a, b, label):
    if isinstance(a, dict):
        assert isinstance(b, dict) and a.keys() == b.keys(), label
        for k in a:
            _equal(a[k], b[k], f"{label}.{k}")
    elif isinstance(a, (tuple, list)):
        assert len(a) == len(b), label
        for i, (x, y) in enumerate(zip(a, b)):
            _equal(x, y, f"{label}[{i}]")
    elif a is None or b is None:
        assert a is None and b is None, label
    else:
        assert np.array_equal(np.asarray(a), np.asarray(b)), f"{label} differs"


def test_autotune_keeps_state():
    import gvom
    import gvom_native as nat
    wl = synth.workload(2, frames=4)
    scans = [st[0] for st in wl.steps]
    dev = [nat.DeviceArray(s.points) for s in scans]
    tuned = gvom.Gvom(*wl.map_args)
    plain = gvom.Gvom(*wl.map_args)

    def frame(g, k):
        s = scans[k % len(scans)]
        g.process_pointcloud_device(dev[k % len(scans)], s.points.shape[0], s.ego + 0.3 * k, s.transform)
        return g.combine_maps()

    for k in range(3):
        frame(tuned, k)
        frame(plain, k)
    before = _snapshot(tuned, tuned.buffer_size)
    s = scans[3]
    block, rpt, flags, best_ms = tuned.autotune(dev[3], s.points.shape[0], s.ego, s.transform)
    assert block in (64, 128, 256) and rpt >= 1 and best_ms > 0
    print(f"autotune chose block {block}, {rpt} rays per thread, flags {flags}, {best_ms:.3f} ms", flush=True)
    _equal(before, _snapshot(tuned, tuned.buffer_size), "state after autotune")
    _equal(before, _snapshot(plain, plain.buffer_size), "state of the untuned handle")
    for k in range(3, 6):
        _equal(frame(tuned, k), frame(plain, k), f"frame {k} maps")
    _equal(_snapshot(tuned, tuned.buffer_size), _snapshot(plain, plain.buffer_size), "state after three more frames")
    print("autotune left the state unchanged; later frames identical", flush=True)
