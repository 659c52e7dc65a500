"""The allocation-free scan entry point on the GPU, through the C-ABI.

gvom_process_pointcloud_device_noalloc must give what gvom_process_pointcloud_device gives
on the same handle state: two calls in a row, the first at the 131,072-point maximum (one
128 x 1024 scan), the second smaller (a 64 x 1024 scan), each followed by combine_maps;
every slot field and every combined / 2D map bit-identical.  A scan past the maximum is
rejected with a clear error and leaves the ring as it was.
"""
import ctypes

import numpy as np
import pytest

import gvom_synth as synth
from oracle.gvom_oracle import canonical_scan

pytestmark = pytest.mark.gpu


def _canon(st):
    return canonical_scan(st["index_map"], st["hit"], st["total"], st["min_h"], st["metrics"])


def _equal_state(a, b, label):
    ca, cb = _canon(a), _canon(b)
    for k in ("occ", "passes", "hit", "total"):
        np.testing.assert_array_equal(ca[k], cb[k], err_msg=f"{label} {k}")
    np.testing.assert_array_equal(ca["min_h"], cb["min_h"], err_msg=f"{label} min_h")
    np.testing.assert_array_equal(ca["metrics"], cb["metrics"], err_msg=f"{label} metrics")
    np.testing.assert_array_equal(a["origin"], b["origin"], err_msg=f"{label} origin")


def test_noalloc_matches_device_entry_point():
    import gvom
    import gvom_native as nat
    head = synth.workload(0, frames=1).steps[0][0]
    small = synth.workload(2, frames=1).steps[0][0]
    assert head.points.shape[0] == 131072 and small.points.shape[0] < 131072
    args = synth.gvom_args(0.4, 0.2, 256, 64, 8)
    ref = gvom.Gvom(*args)
    new = gvom.Gvom(*args)
    lib = new._lib
    new.reserve(131072)   # slot records and the tiled scratch: the calls below allocate nothing
    for k, s in enumerate((head, small)):
        dev = nat.DeviceArray(np.ascontiguousarray(s.points, np.float32))
        n = s.points.shape[0]
        ego = np.ascontiguousarray(s.ego, np.float64)
        T = np.ascontiguousarray(s.transform, np.float32)
        print(f"call {k}: {n} points through gvom_process_pointcloud_device and _noalloc", flush=True)
        nat.check(lib.gvom_process_pointcloud_device(ref._h, dev.ptr, n, nat.GVOM_F32, nat.ptr(ego), nat.ptr(T),
                                                     nat.GVOM_F32), ref._h)
        nat.check(lib.gvom_process_pointcloud_device_noalloc(new._h, dev.ptr, n, nat.GVOM_F32, nat.ptr(ego),
                                                             nat.ptr(T), nat.GVOM_F32), new._h)
        ref.sync()
        new.sync()
        slot = k % 8
        _equal_state(ref.slot_state(slot), new.slot_state(slot), f"call {k} slot {slot}")
        a, b = ref.combine_maps(), new.combine_maps()
        for i, name in enumerate(("origin", "pos", "neg", "rough", "vis")):
            np.testing.assert_array_equal(a[i], b[i], err_msg=f"call {k} {name}")
        ca, cb = ref.combined_state(), new.combined_state()
        _equal_state(ca, cb, f"call {k} combined")
        np.testing.assert_array_equal(canonical_scan(ca["index_map"], ca["hit"], ca["total"], None, ca["eigen"])["metrics"],
                                      canonical_scan(cb["index_map"], cb["hit"], cb["total"], None, cb["eigen"])["metrics"])
        ma, mb = ref.maps_state(), new.maps_state()
        for key in ma:
            np.testing.assert_array_equal(ma[key], mb[key], err_msg=f"call {k} map {key}")
        print(f"call {k}: slot {slot}, combined map and 2D maps equal", flush=True)
    # past the maximum: rejected before any device work, ring unchanged
    big = np.zeros((131073, 3), np.float32)
    big[:, 0] = 5.0
    dev = nat.DeviceArray(big)
    ego = np.zeros(3)
    before = new.ring_state()
    rc = lib.gvom_process_pointcloud_device_noalloc(new._h, dev.ptr, 131073, nat.GVOM_F32, nat.ptr(ego), None,
                                                    nat.GVOM_NO_TRANSFORM)
    assert rc != 0
    msg = lib.gvom_last_error(new._h).decode()
    assert "at most 131072" in msg, msg
    assert new.ring_state() == before
