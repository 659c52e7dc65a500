"""A 1024 x 1024 x 48 map with 30 ring slots through the direct (untiled) ray cast, on the GPU.

The ring's 60 physical slot tables span 3.0e9 words, beyond 2^31: ray keys formed from a
physical buffer index would pass the sign bit that marks a dead lane, and a pass count
would be dropped.  The cast forms keys relative to the scan's place in the batch and reaches
the slot buffer through a 64-bit offset.  A batch of 25 scans fills slots 0-24 (buffers up to
49), a batch of 5 more through the LDS-box variant fills slots 25-29 (buffers up to 59);
slots in both ranges and the combined map are compared with the oracle.
"""
import numpy as np
import pytest

import gvom_synth as synth
from oracle.gvom_oracle import OracleGvom
from parity import assert_close, assert_equal, compare_combined, compare_maps, compare_scan

pytestmark = pytest.mark.gpu


def test_direct_cast_on_a_1024_map_with_30_slots():
    import gvom
    base = synth.workload(1).steps[0][0]
    args = synth.gvom_args(0.4, 0.2, 1024, 48, 30)
    scans = []
    for k in range(30):
        off = np.array([1.3 * k, -0.7 * k, 0.05 * k])
        scans.append(((base.points + off.astype(np.float32)).astype(np.float32), base.ego + off))
    g = gvom.Gvom(*args)
    o = OracleGvom(*args)
    for first, count, box in ((0, 25, False), (25, 5, True)):
        g.set_raycast_config(256, 1, use_box=box, tiled=False)
        part = scans[first:first + count]
        offsets = np.concatenate([[0], np.cumsum([p.shape[0] for p, _ in part])])
        print(f"batch of {count} scans into slots {first}-{first + count - 1} (LDS box {box})", flush=True)
        g.process_batch(np.concatenate([p for p, _ in part]), offsets, np.stack([e for _, e in part]))
        for p, e in part:
            o.Process_pointcloud(p, e)
    for slot in (0, 20, 21, 24, 25, 29):
        compare_scan(g.slot_state(slot), o.slots[slot], f"slot {slot}")
        print(f"slot {slot}: {len(o.slots[slot].hit)} cells equal", flush=True)
    out_g = g.combine_maps()
    out_o = o.combine_maps()
    compare_combined(g.combined_state(), o.combined, "combined")
    compare_maps(g.maps_state(), o.combined, "maps")
    assert_equal("origin_m", out_g[0], out_o[0])
    for name, a, b in zip(("pos", "neg", "vis"), (out_g[1], out_g[2], out_g[4]), (out_o[1], out_o[2], out_o[4])):
        assert_equal(name, a, b)
    assert_close("roughness", out_g[3], out_o[3])
    print(f"combined: {len(o.combined.hit)} cells equal", flush=True)
