"""CPU restatement of the reference's three 2D image kernels that its pipeline never calls --
TEST INFRASTRUCTURE ONLY (imported by tests, never by the product).

Plain Python loops over [x, y] maps (small sizes only), one output cell at a time, in the
order the reference's per-thread bodies run:

- `expand_binary`: Gvom.__expand_binary (gvom.py:807-832).
- `lowpass_binary`: Gvom.__lowpass_binary (gvom.py:834-862).
- `convolve`: Gvom.__convolve (gvom.py:864-885) with its arithmetic as written (every term
  reads input[x + r, y + r] and the bounds test is on x + r, y + r), r = (kernel_size - 1) // 2.
  The reference computes r with true division, a float that Numba's range() rejects, so its
  own kernel never compiles; this restatement therefore has no reference output to match and
  is pinned by closed forms only.
"""
from __future__ import annotations

import math

import numpy as np


def expand_binary(img: np.ndarray, r: float) -> np.ndarray:
    """gvom.py:807-832: 1 where some in-map cell (x + i, y + j) with |i| <= floor(r) and
    |j| <= floor(sqrt(r*r - i*i)) holds exactly 1, else 0."""
    X = img.shape[0]
    out = np.zeros_like(img)
    r_int = int(math.floor(r))
    for x in range(X):
        for y in range(X):
            hit = 0
            for i in range(-r_int, r_int + 1):
                dy = int(math.floor(math.sqrt(r * r - i * i)))                # gvom.py:822
                if x + i < 0 or x + i >= X:
                    continue
                for j in range(-dy, dy + 1):
                    if y + j < 0 or y + j >= X:
                        continue
                    if img[x + i, y + j] == 1:                                 # gvom.py:829
                        hit = 1
                        break
                if hit:
                    break
            out[x, y] = hit
    return out


def lowpass_binary(img: np.ndarray, filter_size: int, filter_fraction: float) -> np.ndarray:
    """gvom.py:834-862: 0 where the input is 0; else 1 when the f64 mean of the input over the
    in-map cells of the (2 f + 1)^2 window reaches filter_fraction."""
    X = img.shape[0]
    out = np.zeros_like(img)
    for x in range(X):
        for y in range(X):
            if img[x, y] == 0:                                                 # gvom.py:847
                continue
            val = 0.0
            cnt = 0.0
            for i in range(-filter_size, filter_size + 1):
                if x + i < 0 or x + i >= X:
                    continue
                for j in range(-filter_size, filter_size + 1):
                    if y + j < 0 or y + j >= X:
                        continue
                    val += float(img[x + i, y + j]) * 1.0                      # gvom.py:857
                    cnt += 1.0
            out[x, y] = 1 if (val / cnt) >= filter_fraction else 0             # gvom.py:860
    return out


def convolve(img: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    """gvom.py:864-885 as written, for an odd kernel size."""
    X = img.shape[0]
    ks = kernel.shape[0]
    r = (ks - 1) // 2
    out = np.zeros(img.shape, np.float64)
    for x in range(X):
        for y in range(X):
            acc = 0.0
            for i in range(-r, r + 1):
                x2 = i + r
                if x + r < 0 or x + r >= X:                                    # gvom.py:877
                    continue
                for j in range(-r, r + 1):
                    y2 = j + r
                    if y + r < 0 or y + r >= X:
                        continue
                    acc += float(img[x + r, y + r]) * float(kernel[x2, y2])    # gvom.py:883
            out[x, y] = acc
    return out
