"""Launch geometry, CPU only: tests/launch_plan_check.cpp sweeps the parameter domain of
gvom_create against the footprint functions of g-vom_amd/csrc/gvom_launch_plan.h (direct and
tiled ray-cast keys below 2^31, slot tables, tiled scratch, column scan, segment counts, radix
histograms), and calls the launchers with buffers one word short, which they must refuse
before any device work.  Built with hipcc as host code against libgvom_hip.so; no GPU."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "g-vom_amd")


def test_launch_plan_and_launcher_refusals():
    lib = os.path.join(PKG, "libgvom_hip.so")
    assert os.path.exists(lib), "build the library first (__graft_entry__.build())"
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "launch_plan_check")
        cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-std=c++17", "-O1", "-I", os.path.join(PKG, "csrc"),
               os.path.join(ROOT, "tests", "launch_plan_check.cpp"), "-L", PKG, "-lgvom_hip",
               f"-Wl,-rpath,{PKG}", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        print(r.stdout)
        assert r.returncode == 0, r.stdout + r.stderr
        assert r.stdout.strip().endswith("ok")
