"""The allocation-free discharge finalisation against gospl_flow_finalize.

gospl_flow_finalize_noalloc takes no mesh handle and allocates nothing; its
results must equal gospl_flow_finalize's bit for bit.  Two calls in a row on
one GPU, the second on a smaller input (the fixed scratch keeps the first
call's partial maxima past the second call's workgroup count).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests.helpers import case

pytestmark = pytest.mark.gpu


def _inputs(n, seed, device):
    import torch

    rng = np.random.default_rng(seed)
    hl = rng.uniform(-100.0, 100.0, n)
    water = hl + np.where(rng.random(n) < 0.3, rng.uniform(0.0, 5.0, n), -1.0)
    pit = np.where(rng.random(n) < 0.4, rng.integers(1, 50, n), -1).astype(np.int32)
    fal = rng.lognormal(0.0, 3.0, n) * np.where(rng.random(n) < 0.05, -1.0, 1.0)
    t = lambda a, dt: torch.as_tensor(a, dtype=dt, device=device)
    return (t(hl, torch.float64), t(water, torch.float64), t(pit, torch.int32), t(fal, torch.float64))


@pytest.mark.parametrize("has_pits", [1, 0])
def test_noalloc_finalize_matches_handle_entry_point(has_pits):
    import torch

    from gospl_amd import _lib
    from gospl_amd.model import DeviceMesh, _ptr

    lib = _lib.load()
    dev = torch.device("cuda:0")
    sizes = []
    for kind, size in (("planar", 120), ("planar", 30)):
        m, _, _ = case(kind, size)
        sizes.append(m)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    for call, m in enumerate(sizes):  # the second call is on a smaller mesh
        dm = DeviceMesh(m, dev)
        n = m.N
        hl, water, pit, fal = _inputs(n, 10 + call, dev)
        fal_a, fal_b = fal.clone(), fal.clone()
        ff_a = torch.empty_like(fal)
        ff_b = torch.empty_like(fal)
        _lib.check(lib.gospl_flow_finalize(dm.handle, _ptr(hl), _ptr(water), _ptr(pit), has_pits, _ptr(fal_a),
                                           _ptr(ff_a), stream), "gospl_flow_finalize")
        _lib.check(lib.gospl_flow_finalize_noalloc(n, _ptr(hl), _ptr(water), _ptr(pit), has_pits, _ptr(fal_b),
                                                   _ptr(ff_b), stream), "gospl_flow_finalize_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(fal_a, fal_b), "call %d (N=%d): fal differs" % (call, n)
        assert torch.equal(ff_a, ff_b), "call %d (N=%d): fillfal differs" % (call, n)
        assert float(ff_b.min()) >= 0.0
        dm.close()


def test_noalloc_finalize_rejects_bad_sizes():
    from gospl_amd import _lib

    lib = _lib.load()
    n = ctypes.c_void_p(0)
    assert lib.gospl_flow_finalize_noalloc(0, n, n, n, 1, n, n, n) != 0
    assert b"outside" in lib.gospl_last_error()
    assert lib.gospl_flow_finalize_noalloc(1 << 31, n, n, n, 1, n, n, n) != 0
