"""The reference's unused 2D image kernels (gvom.py:807-885): the CPU restatement pinned by
closed forms and hand-worked cases (no GPU needed)."""
import numpy as np

from oracle import image_ref


def test_expand_radius_zero_is_identity():
    rng = np.random.default_rng(1)
    img = (rng.random((13, 13)) < 0.3).astype(np.int32)
    assert np.array_equal(image_ref.expand_binary(img, 0.0), img)
    assert np.array_equal(image_ref.expand_binary(img, 0.99), img)


def test_expand_single_pixel_gives_the_disk():
    # r = 2.5: rows i = 0, +-1 reach |j| <= floor(sqrt(6.25 - i^2)) = 2, rows i = +-2 reach 1
    img = np.zeros((9, 9), np.int32)
    img[4, 4] = 1
    want = np.zeros((9, 9), np.int32)
    want[3:6, 2:7] = 1
    want[2, 3:6] = 1
    want[6, 3:6] = 1
    assert np.array_equal(image_ref.expand_binary(img, 2.5), want)
    # the disk is clipped, not wrapped, at the map edge
    img = np.zeros((9, 9), np.int32)
    img[0, 8] = 1
    out = image_ref.expand_binary(img, 1.0)
    assert out.sum() == 3 and out[0, 8] == out[1, 8] == out[0, 7] == 1


def test_expand_only_counts_ones():
    img = np.full((5, 5), 2, np.int32)   # non-binary values are not "1"
    assert image_ref.expand_binary(img, 3.0).sum() == 0


def test_lowpass_hand_worked_block():
    img = np.zeros((7, 7), np.int32)
    img[2:5, 2:5] = 1
    out = image_ref.lowpass_binary(img, 1, 0.5)
    want = np.zeros((7, 7), np.int32)
    want[3, 3] = 1                                  # 9 / 9
    want[2, 3] = want[4, 3] = want[3, 2] = want[3, 4] = 1   # edge middles: 6 / 9
    assert np.array_equal(out, want)                # corners: 4 / 9 < 0.5


def test_lowpass_fraction_limits_and_edges():
    rng = np.random.default_rng(2)
    img = (rng.random((11, 11)) < 0.5).astype(np.int32)
    assert np.array_equal(image_ref.lowpass_binary(img, 2, 0.0), img)
    assert image_ref.lowpass_binary(img, 2, 1.01).sum() == 0
    # a corner cell's window holds only its 4 in-map cells
    img = np.zeros((6, 6), np.int32)
    img[0, 0] = img[0, 1] = img[1, 0] = 1
    assert image_ref.lowpass_binary(img, 1, 0.75)[0, 0] == 1   # 3 / 4
    assert image_ref.lowpass_binary(img, 1, 0.76)[0, 0] == 0


def test_convolve_as_written():
    rng = np.random.default_rng(3)
    img = rng.random((8, 8))
    k = rng.random((3, 3))
    out = image_ref.convolve(img, k)
    # every term reads img[x + 1, y + 1], so a cell is that value times the kernel's sum;
    # cells whose (x + 1, y + 1) leaves the map (last row and column) are 0
    want = np.zeros((8, 8))
    want[:7, :7] = img[1:, 1:] * k.sum()
    assert np.allclose(out, want, rtol=1e-14, atol=0)
    assert np.all(out[7, :] == 0) and np.all(out[:, 7] == 0)
    # a 1 x 1 kernel scales the image; a zero kernel gives zeros
    assert np.array_equal(image_ref.convolve(img, np.array([[2.0]])), img * 2.0)
    assert not image_ref.convolve(img, np.zeros((5, 5))).any()
