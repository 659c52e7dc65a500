"""The reference's unused 2D image kernels (gvom.py:807-885) on an MI355X against their CPU
restatement (oracle/image_ref.py)."""
import numpy as np
import pytest

import gvom_synth as synth
from parity import assert_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size", [64, 37])
def test_image_kernels_match_oracle(size):
    """The reference's unused image kernels (gvom.py:807-885) on the device against their CPU
    restatement: random 0/1 maps, several radii, window sizes and fractions, and the
    convolution on a random f64 map; bit-exact."""
    import gvom
    from oracle import image_ref
    wl = synth.workload(1)
    args = (0.4, 0.2, size, 16, 1) + tuple(wl.map_args[5:])
    g = gvom.Gvom(*args)
    rng = np.random.default_rng(size)
    img = (rng.random((size, size)) < 0.08).astype(np.int32)
    for r in (0.0, 1.0, 1.5, 2.9, 4.0):
        assert_equal(f"expand r={r}", g.expand_binary(img, r), image_ref.expand_binary(img, r))
    dense = (rng.random((size, size)) < 0.6).astype(np.int32)
    for f, frac in ((1, 0.5), (2, 0.7), (0, 1.0), (3, 0.0)):
        assert_equal(f"lowpass f={f} frac={frac}", g.lowpass_binary(dense, f, frac),
                     image_ref.lowpass_binary(dense, f, frac))
    fimg = rng.normal(size=(size, size))
    for ks in (1, 3, 5):
        k = rng.normal(size=(ks, ks))
        assert_equal(f"convolve k={ks}", g.convolve(fimg, k), image_ref.convolve(fimg, k))
    print(f"image kernels {size}: ok", flush=True)
