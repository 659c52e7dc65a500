"""Batched scans and debug exporters on the GPU.

- process_batch over several sensors' ragged scans gives bit-identical slots, combined
  maps and 2D maps to the same scans sent one Process_pointcloud call at a time, and
  matches the oracle.
- make_debug_voxel_map / make_debug_height_map / make_debug_inferred_height_map match
  the oracle's restatement of gvom.py:395-503 (rows compared after ordering by voxel).
"""
import numpy as np
import pytest

import gvom_synth as synth
from oracle.gvom_oracle import OracleGvom
from parity import assert_close, assert_equal, compare_combined, compare_maps, compare_scan

pytestmark = pytest.mark.gpu


def _state_equal(a, b, label):
    for k in a:
        assert_equal(f"{label} {k}", a[k], b[k])


@pytest.fixture(scope="module")
def cfg4():
    return synth.workload(4, frames=3)


def test_batch_equals_sequential_and_oracle(cfg4):
    import gvom
    wl = cfg4
    seq = gvom.Gvom(*wl.map_args)
    bat = gvom.Gvom(*wl.map_args)
    orc = OracleGvom(*wl.map_args)
    for k, step in enumerate(wl.steps):
        for s in step:
            seq.Process_pointcloud(s.points, s.ego, s.transform)
            orc.Process_pointcloud(s.points, s.ego, s.transform)
        pts = np.concatenate([s.points for s in step])
        off = np.cumsum([0] + [s.points.shape[0] for s in step])
        bat.process_batch(pts, off, np.stack([s.ego for s in step]), np.stack([s.transform for s in step]))
        assert seq.ring_state() == bat.ring_state()
        for slot in range(wl.map_args[4]):
            a, b = seq.slot_state(slot), bat.slot_state(slot)
            assert (a is None) == (b is None)
            if a is not None:
                _state_equal(a, b, f"step {k} slot {slot}")
                compare_scan(b, orc.slots[slot], f"step {k} slot {slot} vs oracle")
        out_s = seq.combine_maps()
        out_b = bat.combine_maps()
        orc.combine_maps()
        for x, y in zip(out_s, out_b):
            assert_equal(f"step {k} outputs", x, y)
        _state_equal(seq.combined_state(), bat.combined_state(), f"step {k} combined")
        _state_equal(seq.maps_state(), bat.maps_state(), f"step {k} maps")
        compare_combined(bat.combined_state(), orc.combined, f"step {k} combined vs oracle")
        compare_maps(bat.maps_state(), orc.combined, f"step {k} maps vs oracle")
        print(f"batched step {k}: {len(step)} scans, {sum(s.points.shape[0] for s in step)} points ok", flush=True)


def test_cfg4_all_eight_steps_against_the_oracle():
    """Configuration 4 at full length: 3 sensors x 8 steps (24 scans of 128/64/16 x 1024 rays,
    10-40 % dropout) through process_batch, one combine per step, buffer 8: the ring fills
    after the third step and wraps twice (gvom.py:198-216, 242-298).  The 3 slots each step
    writes, the combined map and the 2D maps are compared with the oracle every step."""
    import gvom
    wl = synth.workload(4, frames=8)
    bat = gvom.Gvom(*wl.map_args)
    orc = OracleGvom(*wl.map_args)
    for k, step in enumerate(wl.steps):
        written = []
        for s in step:
            orc.Process_pointcloud(s.points, s.ego, s.transform)
            written.append(orc.last_buffer_index)
        pts = np.concatenate([s.points for s in step])
        off = np.cumsum([0] + [s.points.shape[0] for s in step])
        bat.process_batch(pts, off, np.stack([s.ego for s in step]), np.stack([s.transform for s in step]))
        for slot in written:
            compare_scan(bat.slot_state(slot), orc.slots[slot], f"step {k} slot {slot}")
        out_b = bat.combine_maps()
        out_o = orc.combine_maps()
        compare_combined(bat.combined_state(), orc.combined, f"step {k} combined")
        compare_maps(bat.maps_state(), orc.combined, f"step {k} maps")
        for name, a, b in zip(("origin", "pos", "neg", "vis"), (out_b[0], out_b[1], out_b[2], out_b[4]),
                              (out_o[0], out_o[1], out_o[2], out_o[4])):
            assert_equal(f"step {k} {name}", a, b)
        print(f"cfg4 step {k}: slots {written}, {len(orc.combined.hit)} combined cells ok", flush=True)


def _sort_rows(a, keycols):
    order = np.lexsort(tuple(a[:, c] for c in reversed(keycols)))
    return a[order]


def test_debug_exporters_match_oracle():
    import gvom
    wl = synth.workload(2, frames=2)
    g = gvom.Gvom(*wl.map_args)
    o = OracleGvom(*wl.map_args)
    assert g.make_debug_height_map() is None
    for step in wl.steps:
        for s in step:
            g.Process_pointcloud(s.points, s.ego, s.transform)
            o.Process_pointcloud(s.points, s.ego, s.transform)
        g.combine_maps()
        o.combine_maps()
    vg = _sort_rows(g.make_debug_voxel_map(), [0, 1, 2])
    vo = _sort_rows(o.make_debug_voxel_map(), [0, 1, 2])
    assert vg.shape == vo.shape
    assert_equal("voxel xyz / solid factor / count", vg[:, :5], vo[:, :5])
    tr = np.abs(vo[:, 5]) + np.abs(vo[:, 6]) + 3 * np.abs(vo[:, 7]) + 1e-12
    ok = np.isclose(vg[:, 5:], vo[:, 5:], rtol=1e-4, atol=1e-5) | (np.abs(vg[:, 5:] - vo[:, 5:]) <= 1e-5 * tr[:, None])
    assert ok.all(), f"{(~ok).sum()} eigen-gap entries differ"
    hg, ho = g.make_debug_height_map(), o.make_debug_height_map()
    assert_equal("height rows xyz", hg[:, :3], ho[:, :3])
    assert_close("height rows slopes/roughness", hg[:, 3:], ho[:, 3:])
    ig, io = g.make_debug_inferred_height_map(), o.make_debug_inferred_height_map()
    assert_equal("inferred rows", ig, io)


def test_graph_replay_matches_eager(cfg4):
    """A captured sequence (2 steps of 3-scan batch + combine) replayed three times, with
    eager calls in between, leaves every state bit-identical to running the same calls
    eagerly."""
    import gvom
    import gvom_native as nat
    wl = cfg4
    steps = wl.steps[:2]
    flat = [np.concatenate([s.points for s in st]) for st in steps]
    offs = [np.cumsum([0] + [s.points.shape[0] for s in st]) for st in steps]
    egos = [np.stack([s.ego for s in st]) for st in steps]
    Ts = [np.stack([s.transform for s in st]) for st in steps]
    dev = [nat.DeviceArray(f) for f in flat]
    eager = gvom.Gvom(*wl.map_args)
    graph_h = gvom.Gvom(*wl.map_args)
    for g in (eager, graph_h):
        g.reserve(max(s.points.shape[0] for st in steps for s in st), max(f.shape[0] for f in flat))
    # one eager step first so that capture starts from a non-empty ring
    for g in (eager, graph_h):
        g.process_batch_device(dev[0], offs[0], egos[0], Ts[0])
        g.combine_maps_device()
    graph_h.sync()
    graph_h.graph_begin(16)
    for k in range(2):
        graph_h.process_batch_device(dev[k], offs[k], egos[k], Ts[k])
        graph_h.combine_maps_device()
    gr = graph_h.graph_end()
    assert gr.scans == 6
    shift = np.array([0.4, 0.0, 0.0])
    for r in range(3):
        e_all = np.concatenate([egos[0] + r * shift, egos[1] + r * shift])
        gr.launch(e_all, np.concatenate(Ts))
        for k in range(2):
            eager.process_batch_device(dev[k], offs[k], egos[k] + r * shift, Ts[k])
            eager.combine_maps_device()
        a, b = eager.combine_maps(), graph_h.combine_maps()     # eager work between replays
        for x, y in zip(a, b):
            assert_equal(f"replay {r} outputs", x, y)
        assert eager.ring_state() == graph_h.ring_state()
        _state_equal(eager.combined_state(), graph_h.combined_state(), f"replay {r} combined")
        for slot in range(wl.map_args[4]):
            a_s, b_s = eager.slot_state(slot), graph_h.slot_state(slot)
            assert (a_s is None) == (b_s is None)
            if a_s is not None:
                _state_equal(a_s, b_s, f"replay {r} slot {slot}")
        print(f"graph replay {r} ok", flush=True)


@pytest.mark.parametrize("flags", [4, 12, 6, 1, 9, 2, 3, 8])
def test_ray_cast_variants_equal(cfg4, flags):
    """Every ray-cast variant (tiled scratch, sector order, LDS box, spread wavefronts) leaves
    slot tables and records bit-identical to the plain cast, for single scans and ragged
    3-scan batches, and so do the combined map and the 2D maps."""
    import gvom
    wl = cfg4
    a = gvom.Gvom(*wl.map_args)
    a.set_raycast_config(256, 1, False, False, False, False)
    b = gvom.Gvom(*wl.map_args)
    b.set_raycast_config(128, 2, bool(flags & 1), bool(flags & 2), bool(flags & 4), bool(flags & 8))
    for k, step in enumerate(wl.steps):
        if k == 0:
            for s in step:
                a.Process_pointcloud(s.points, s.ego, s.transform)
                b.Process_pointcloud(s.points, s.ego, s.transform)
        else:
            pts = np.concatenate([s.points for s in step])
            off = np.cumsum([0] + [s.points.shape[0] for s in step])
            egos = np.stack([s.ego for s in step])
            Ts = np.stack([s.transform for s in step])
            a.process_batch(pts, off, egos, Ts)
            b.process_batch(pts, off, egos, Ts)
        for slot in range(wl.map_args[4]):
            sa, sb = a.slot_state(slot), b.slot_state(slot)
            assert (sa is None) == (sb is None)
            if sa is not None:
                _state_equal(sa, sb, f"flags {flags} step {k} slot {slot}")
        a.combine_maps()
        b.combine_maps()
        _state_equal(a.combined_state(), b.combined_state(), f"flags {flags} step {k} combined")
        _state_equal(a.maps_state(), b.maps_state(), f"flags {flags} step {k} maps")
        print(f"ray variant {flags}: step {k} slots and combined map identical", flush=True)


@pytest.mark.parametrize("rings", [(1024, 32, 8), (1024, 16, 12), (1000, 8, 8), (1024, 64, 0), (960, 2, 8)])
def test_ring_patch_ray_order_equal(cfg4, rings):
    """The ring-patch ray order (gvom_set_ray_rings: patches of 64 / pr azimuths x pr rings)
    leaves slot tables, records, the combined map and the 2D maps bit-identical to the linear
    order, for single scans and ragged 3-scan batches (whose rings are 1024 points before
    dropout, so the patches only approximate rings), with the direct and the tiled cast, and
    for strides that do not divide the scans."""
    import gvom
    stride, pr, flags = rings
    wl = cfg4
    a = gvom.Gvom(*wl.map_args)
    a.set_raycast_config(256, 1, False, False, False, False)
    b = gvom.Gvom(*wl.map_args)
    b.set_raycast_config(256, 1, bool(flags & 1), bool(flags & 2), bool(flags & 4), bool(flags & 8))
    b.set_ray_rings(stride, pr)
    for k, step in enumerate(wl.steps):
        if k == 0:
            for s in step:
                a.Process_pointcloud(s.points, s.ego, s.transform)
                b.Process_pointcloud(s.points, s.ego, s.transform)
        else:
            pts = np.concatenate([s.points for s in step])
            off = np.cumsum([0] + [s.points.shape[0] for s in step])
            egos = np.stack([s.ego for s in step])
            Ts = np.stack([s.transform for s in step])
            a.process_batch(pts, off, egos, Ts)
            b.process_batch(pts, off, egos, Ts)
        for slot in range(wl.map_args[4]):
            sa, sb = a.slot_state(slot), b.slot_state(slot)
            assert (sa is None) == (sb is None)
            if sa is not None:
                _state_equal(sa, sb, f"rings {rings} step {k} slot {slot}")
        a.combine_maps()
        b.combine_maps()
        _state_equal(a.combined_state(), b.combined_state(), f"rings {rings} step {k} combined")
        _state_equal(a.maps_state(), b.maps_state(), f"rings {rings} step {k} maps")
        print(f"ring patches {rings}: step {k} slots and combined map identical", flush=True)


def test_ring_patch_hint_rejected():
    import gvom
    import gvom_native as nat
    g = gvom.Gvom(*synth.gvom_args(0.4, 0.2, 64, 32, 2))
    for stride, pr in ((1023, 32), (1024, 3), (1024, 128), (1, 32)):
        with pytest.raises(nat.GvomError):
            g.set_ray_rings(stride, pr)
    g.set_ray_rings(0, 0)


@pytest.mark.parametrize("variant", [(128, 1, 4), (256, 1, 0)])
def test_overlap_matches_serial(variant):
    """With scan/combine overlap on, device-pointer frames issued back to back (eager, then
    a captured 4-frame sequence replayed twice, then two-scan batches) leave the ring, every
    slot, the combined map and the 2D maps bit-identical to the same frames without overlap;
    tiled and direct ray casts."""
    import gvom
    import gvom_native as nat
    wl = synth.workload(2)
    scans = [st[0] for st in wl.steps]
    dev = [nat.DeviceArray(s.points) for s in scans]
    serial = gvom.Gvom(*wl.map_args)
    ovl = gvom.Gvom(*wl.map_args)
    for g in (serial, ovl):
        g.set_raycast_config(*variant[:2], False, False, bool(variant[2] & 4))
    ovl.set_overlap(True)

    def frame(g, k, shift):
        s = scans[k % len(scans)]
        g.process_pointcloud_device(dev[k % len(scans)], s.points.shape[0], s.ego + shift, s.transform)
        g.combine_maps_device()

    def compare(label):
        assert serial.ring_state() == ovl.ring_state()
        _state_equal(serial.combined_state(), ovl.combined_state(), f"{label} combined")
        _state_equal(serial.maps_state(), ovl.maps_state(), f"{label} maps")
        for slot in range(wl.map_args[4]):
            a_s, b_s = serial.slot_state(slot), ovl.slot_state(slot)
            assert (a_s is None) == (b_s is None)
            if a_s is not None:
                _state_equal(a_s, b_s, f"{label} slot {slot}")

    zero = np.zeros(3)
    for k in range(6):
        frame(serial, k, zero)
        frame(ovl, k, zero)
    compare("eager")
    print("overlap eager: 6 frames identical", flush=True)
    ovl.sync()
    ovl.graph_begin(16)
    for k in range(4):
        frame(ovl, k, zero)
    gr = ovl.graph_end()
    shift = np.array([0.4, -0.2, 0.1])
    for r in range(2):
        egos = np.stack([scans[k].ego + r * shift for k in range(4)])
        gr.launch(egos, np.stack([scans[k].transform for k in range(4)]))
        for k in range(4):
            frame(serial, k, r * shift)
        compare(f"replay {r}")
        print(f"overlap graph replay {r}: identical", flush=True)
    # two-scan batches back to back
    flat = nat.DeviceArray(np.concatenate([scans[0].points, scans[1].points]))
    off = [0, scans[0].points.shape[0], scans[0].points.shape[0] + scans[1].points.shape[0]]
    egos = np.stack([scans[0].ego, scans[1].ego])
    Ts = np.stack([scans[0].transform, scans[1].transform])
    for k in range(3):
        for g in (serial, ovl):
            g.process_batch_device(flat, off, egos + k * shift, Ts)
            g.combine_maps_device()
    compare("batches")
    print("overlap batches: identical", flush=True)


@pytest.mark.parametrize("cfg", [1, 2])
def test_uneven_cadence_overlap_serial_oracle(cfg):
    """Scans and combines in an uneven order (two scans per combine, a combine with no new
    scan, three scans in a row), the sensor moving, with overlap on and off, each combine
    checked against the oracle.  Buffer 1 (cfg1: every scan stage writes the one slot's
    other buffer while the previous combine may still read it) and buffer 4 (cfg2, the ring
    wrapping)."""
    import gvom
    import gvom_native as nat
    wl = synth.workload(cfg)
    scans = [st[0] for st in wl.steps]
    dev = [nat.DeviceArray(s.points) for s in scans]
    serial = gvom.Gvom(*wl.map_args)
    ovl = gvom.Gvom(*wl.map_args)
    ovl.set_overlap(True)
    o = OracleGvom(*wl.map_args)
    shift = np.array([0.45, -0.3, 0.2])
    plan = "PPCCPCCPPPCPC"
    k = 0
    for step, op in enumerate(plan):
        label = f"cfg{cfg} op {step} ({op})"
        if op == "P":
            s = scans[k % len(scans)]
            ego = s.ego + k * shift
            for g in (serial, ovl):
                g.process_pointcloud_device(dev[k % len(scans)], s.points.shape[0], ego, s.transform)
            o.Process_pointcloud(s.points, ego, s.transform)
            k += 1
            continue
        origins = [g.combine_maps_device() for g in (serial, ovl)]
        out_o = o.combine_maps()
        assert out_o is not None
        for g, mode, org in ((serial, "serial", origins[0]), (ovl, "overlap", origins[1])):
            assert_equal(f"{label} {mode} origin_m", org, out_o[0])
            assert g.ring_state()["buffer_index"] == o.buffer_index
            assert g.ring_state()["last_buffer_index"] == o.last_buffer_index
            slot = o.last_buffer_index
            compare_scan(g.slot_state(slot), o.slots[slot], f"{label} {mode} slot {slot}")
            compare_combined(g.combined_state(), o.combined, f"{label} {mode} combined")
            compare_maps(g.maps_state(), o.combined, f"{label} {mode}")
        print(f"{label}: cells {len(o.combined.hit)} ok", flush=True)


def test_largest_ring_and_batch():
    """The largest ring (buffer 30) filled by one 30-scan batch, then a 7-scan batch that
    wraps it, each followed by a combine: slots, combined map and 2D maps bit-identical to
    the same scans sent one call at a time, and equal to the oracle (64 x 64 x 32 map, the
    16 x 512 scan with the ego moved per scan)."""
    import gvom
    wl = synth.workload(1)
    s0 = wl.steps[0][0]
    args = synth.gvom_args(0.4, 0.2, 64, 32, 30)
    rng = np.random.default_rng(30)
    egos = [s0.ego + np.array([rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), rng.uniform(-0.5, 0.5)])
            for _ in range(37)]
    seq = gvom.Gvom(*args)
    bat = gvom.Gvom(*args)
    o = OracleGvom(*args)
    for lo, hi in ((0, 30), (30, 37)):
        n = hi - lo
        pts = np.concatenate([s0.points] * n)
        off = np.arange(n + 1) * s0.points.shape[0]
        bat.process_batch(pts, off, np.stack(egos[lo:hi]), np.stack([s0.transform] * n))
        for k in range(lo, hi):
            seq.Process_pointcloud(s0.points, egos[k], s0.transform)
            o.Process_pointcloud(s0.points, egos[k], s0.transform)
        assert bat.ring_state() == seq.ring_state()
        for slot in range(30):
            a, b = seq.slot_state(slot), bat.slot_state(slot)
            assert (a is None) == (b is None)
            if a is not None:
                _state_equal(a, b, f"batch {lo}-{hi} slot {slot}")
        compare_scan(bat.slot_state(o.last_buffer_index), o.slots[o.last_buffer_index], f"batch {lo}-{hi} newest")
        out_b = bat.combine_maps()
        seq.combine_maps()
        out_o = o.combine_maps()
        _state_equal(seq.combined_state(), bat.combined_state(), f"batch {lo}-{hi} combined")
        _state_equal(seq.maps_state(), bat.maps_state(), f"batch {lo}-{hi} maps")
        compare_combined(bat.combined_state(), o.combined, f"batch {lo}-{hi} combined vs oracle")
        compare_maps(bat.maps_state(), o.combined, f"batch {lo}-{hi}")
        for name, a, b in zip(("pos", "neg", "vis"), (out_b[1], out_b[2], out_b[4]), (out_o[1], out_o[2], out_o[4])):
            assert_equal(f"batch {lo}-{hi} {name}", a, b)
        print(f"ring 30, batch {lo}-{hi}: cells {len(o.combined.hit)} ok", flush=True)
