import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "g-vom_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

# Load libgvom_hip.so before any test module imports torch.  torch's bundled HIP runtime has
# the same SONAME (libamdhip64.so.7) as the ROCm runtime the library is built against; loaded
# first, it would serve the library too, and the GPU tests would run on another runtime.
try:
    import gvom_native

    gvom_native.load()
except Exception:   # not built yet: the tests that need it fail on their own
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
