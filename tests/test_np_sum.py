"""The native flow driver sums the cascade's spilled volumes as numpy does
(its decisions compare that sum with thresholds): gospl_np_sum equals
np.sum bit for bit on arrays of every length class (< 8, <= 128, split
recursively above, and past numpy's 8192-element reduction buffer) and
magnitudes spanning 11 decades.  CPU only (a host
function of the library)."""

import ctypes

import numpy as np

from gospl_amd import _lib


def test_np_sum_matches_numpy_bit_for_bit():
    lib = _lib.load()
    rng = np.random.default_rng(3)
    for n in list(range(0, 300)) + [511, 512, 513, 1000, 4096, 8192, 8193, 10007, 20000, 40001]:
        a = rng.normal(size=n) * 10.0 ** rng.integers(-3, 8, size=n)
        a = np.ascontiguousarray(a)
        got = lib.gospl_np_sum(a.ctypes.data_as(ctypes.c_void_p), n)
        assert got == float(np.sum(a)), n
