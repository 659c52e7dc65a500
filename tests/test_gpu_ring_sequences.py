"""Ring sequences beyond one scan per combine (runs on an MI355X).

Every combined map and 2D map is compared with the oracle along a sequence that uses the
ring in every way the API allows: a map of 96 z levels (two 64-word chunks per column in the
row combine, the second partial) whose origin moves up, down and sideways between frames;
two scans processed before one combine; a combine repeated without a new scan; a batch of
three scans; a ring of four slots that wraps.
"""
import numpy as np
import pytest

import gvom_synth as synth
from oracle.gvom_oracle import OracleGvom
from parity import compare_combined, compare_maps, assert_equal

pytestmark = pytest.mark.gpu

DZ = [0.0, 0.55, -0.75, 1.35, -2.15, 0.25, 3.05, -0.45, 0.95, -1.25]


def _scan(k, base):
    return synth.ScanInput(base.points, base.ego + np.array([0.45 * k, -0.3 * k, DZ[k]]), base.transform, k)


def _check(g, o, label):
    out_g = g.combine_maps()
    out_o = o.combine_maps()
    compare_combined(g.combined_state(), o.combined, f"{label} combined")
    compare_maps(g.maps_state(), o.combined, label)
    assert_equal(f"{label} origin_m", out_g[0], out_o[0])
    for name, a, b in zip(("pos", "neg", "vis"), (out_g[1], out_g[2], out_g[4]), (out_o[1], out_o[2], out_o[4])):
        assert_equal(f"{label} {name}", a, b)
    print(f"{label}: cells {len(o.combined.hit)} ok", flush=True)


@pytest.mark.parametrize("variant", [(256, 1, 0), (128, 1, 4)])
def test_ring_sequences(variant):
    import gvom
    wl = synth.workload(1)
    base = wl.steps[0][0]
    args = (0.4, 0.1, 64, 96, 4) + tuple(wl.map_args[5:])
    g = gvom.Gvom(*args)
    g.set_raycast_config(*variant[:2], False, False, bool(variant[2] & 4))
    o = OracleGvom(*args)
    scans = [_scan(k, base) for k in range(len(DZ))]

    def proc(s):
        g.Process_pointcloud(s.points, s.ego, s.transform)
        o.Process_pointcloud(s.points, s.ego, s.transform)

    proc(scans[0])
    _check(g, o, f"ring {variant} frame 0")
    proc(scans[1])
    _check(g, o, f"ring {variant} frame 1")
    # two scans, one combine
    proc(scans[2])
    proc(scans[3])
    _check(g, o, f"ring {variant} frames 2+3")
    # a second combine of the same ring
    _check(g, o, f"ring {variant} repeat")
    proc(scans[4])
    _check(g, o, f"ring {variant} frame 4 (ring wrapped)")
    # three scans in one batch, then one combine
    batch = scans[5:8]
    pts = np.concatenate([s.points for s in batch])
    off = np.cumsum([0] + [len(s.points) for s in batch])
    g.process_batch(pts, off, np.stack([s.ego for s in batch]), np.stack([s.transform for s in batch]))
    for s in batch:
        o.Process_pointcloud(s.points, s.ego, s.transform)
    _check(g, o, f"ring {variant} batch 5-7")
    for k in (8, 9):
        proc(scans[k])
        _check(g, o, f"ring {variant} frame {k}")
