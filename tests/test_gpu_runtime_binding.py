"""The library bound to torch's bundled HIP runtime (torch imported first, in a fresh process).

Both runtimes carry the SONAME libamdhip64.so.7, so a process that imports torch before gvom
runs libgvom_hip.so on torch's copy (older than the ROCm runtime the library is built
against on this image).  Under an older runtime a captured sequence does not fork the
overlapped scan stage onto further streams; this test runs eager overlapped frames and a
captured, overlapped 4-frame sequence replayed twice in that binding, and checks the ring,
every slot, the combined map and the 2D maps bit-identical to serial eager frames.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import torch  # noqa: F401  (first: its HIP runtime then serves libgvom_hip.so too)
sys.path[:0] = [{pkg!r}, {root!r}]
import numpy as np
import gvom
import gvom_native as nat
import gvom_synth as synth

built, running = nat.hip_versions()
print("HIP built", built, "running", running, "runtimes", nat.hip_runtime_files(), flush=True)
wl = synth.workload(2)
scans = [st[0] for st in wl.steps]
dev = [nat.DeviceArray(s.points) for s in scans]
serial = gvom.Gvom(*wl.map_args)
ovl = gvom.Gvom(*wl.map_args)
ovl.set_overlap(True)


def frame(g, k, shift):
    s = scans[k % len(scans)]
    g.process_pointcloud_device(dev[k % len(scans)], s.points.shape[0], s.ego + shift, s.transform)
    g.combine_maps_device()


def same(a, b, label):
    for key in a:
        assert np.array_equal(a[key], b[key]), (label, key)


def compare(label):
    assert serial.ring_state() == ovl.ring_state(), label
    same(serial.combined_state(), ovl.combined_state(), label + " combined")
    same(serial.maps_state(), ovl.maps_state(), label + " maps")
    for slot in range(wl.map_args[4]):
        a, b = serial.slot_state(slot), ovl.slot_state(slot)
        assert (a is None) == (b is None), (label, slot)
        if a is not None:
            same(a, b, label + " slot %d" % slot)


zero = np.zeros(3)
for k in range(6):
    frame(serial, k, zero)
    frame(ovl, k, zero)
compare("eager")
print("eager overlapped frames identical", flush=True)
ovl.sync()
ovl.graph_begin(16)
for k in range(4):
    frame(ovl, k, zero)
gr = ovl.graph_end()
print("captured", flush=True)
shift = np.array([0.4, -0.2, 0.1])
for r in range(2):
    egos = np.stack([scans[k].ego + r * shift for k in range(4)])
    gr.launch(egos, np.stack([scans[k].transform for k in range(4)]))
    for k in range(4):
        frame(serial, k, r * shift)
    compare("replay %d" % r)
    print("replay", r, "identical", flush=True)
print("ok", flush=True)
"""


def test_torch_runtime_binding_overlap_and_capture():
    code = CHILD.format(pkg=os.path.join(ROOT, "g-vom_amd"), root=ROOT)
    print("child: torch imported first, overlap on, eager + captured frames", flush=True)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert r.stdout.strip().endswith("ok")


CHILD_AUTOGRAD = r"""
import sys
import torch
sys.path[:0] = [{pkg!r}, {root!r}]
import numpy as np
import gvom_native as nat
import gvom_synth as synth
from gvom_torch import MapLoss, map_loss

print("runtimes", nat.hip_runtime_files(), flush=True)
s = synth.workload(5).steps[0][0]
T0 = synth.perturb_transform(s.transform, np.random.default_rng(105)).astype(np.float64)
pts = s.points.astype(np.float64)
# autograd path: float64 CUDA points, 6-vector xi at 0
p = torch.tensor(pts, device="cuda:0", requires_grad=True)
xi = torch.zeros(6, dtype=torch.float64, requires_grad=True)
L = map_loss(p, xi, T0, s.ego, MapLoss(0.4, 0.2, 256, 64))
L.backward()
gp = p.grad.detach().cpu().numpy()
gx = xi.grad.numpy()
# the same kernels through MapLoss directly on a DeviceArray
ml = MapLoss(0.4, 0.2, 256, 64)
dev = nat.DeviceArray(pts)
gdev = nat.DeviceArray(np.zeros_like(pts))
out = ml.forward(dev.ptr.value, pts.shape[0], T0, s.ego)
gxi = ml.backward(1.0, gdev.ptr.value)
g_direct = gdev.to_host()
print("L", float(L), out[0], flush=True)
assert L.is_cuda and float(L) == out[0]
np.testing.assert_array_equal(gx, gxi)
np.testing.assert_array_equal(gp, g_direct)
assert np.abs(gp).max() > 0
# xi on the device and an upstream gradient of 2: on a side stream whose points come from a
# kernel just enqueued there (the library's work must be ordered after it and before the
# gradient reads), without host synchronisation inside map_loss
ml2 = MapLoss(0.4, 0.2, 256, 64)
s1 = torch.cuda.Stream()
base = torch.tensor(pts, device="cuda:0")
torch.cuda.synchronize()
with torch.cuda.stream(s1):
    p3 = (base * 1.0).requires_grad_()
    xi3 = torch.zeros(6, dtype=torch.float64, device="cuda:0", requires_grad=True)
    L3 = map_loss(p3, xi3, T0, s.ego, ml2)
    (2.0 * L3).backward()
    gp3 = p3.grad * 1.0
    gx3 = xi3.grad * 1.0
s1.synchronize()
gxi2 = ml.backward(2.0, gdev.ptr.value)
assert xi3.grad.is_cuda
assert float(L3) == out[0]
np.testing.assert_array_equal(gx3.cpu().numpy(), gxi2)
np.testing.assert_array_equal(gp3.cpu().numpy(), gdev.to_host())
print("ok", flush=True)
"""


def test_map_loss_autograd_with_torch_cuda_tensors():
    """torch imported first (one HIP runtime for torch and the library): map_loss over a
    float64 CUDA tensor, on torch's current stream through the stream-ordered entry points,
    gives the loss, the point gradient (a CUDA tensor) and the se(3) gradient of direct
    synchronous MapLoss forward/backward calls on the same scan, bit for bit; also with xi on
    the device, an upstream gradient of 2 and a non-default stream."""
    code = CHILD_AUTOGRAD.format(pkg=os.path.join(ROOT, "g-vom_amd"), root=ROOT)
    print("child: torch imported first, map_loss on cuda:0", flush=True)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert r.stdout.strip().endswith("ok")
