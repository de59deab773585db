"""The allocation-free branching entry point (jqmc_lrdmc_branching_static) against
jqmc_lrdmc_branching through the C ABI: identical parent indices and weighted sums (the same comb
kernel, bit for bit) on two calls in a row, the second with fewer walkers than the first, and a walker
count above the static maximum refused before any launch."""

import ctypes

import numpy as np
import pytest
import torch

from jqmc_amd import _lib

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _both(W, seed):
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda", 0)
    w = torch.as_tensor(rng.uniform(0.2, 2.0, W), dtype=torch.float64, device=dev)
    vd = torch.as_tensor(rng.uniform(20.0, 30.0, W), dtype=torch.float64, device=dev)
    eL = torch.as_tensor(rng.normal(-15.0, 1.0, W), dtype=torch.float64, device=dev)
    zeta = float(rng.random())
    stream = torch.cuda.current_stream().cuda_stream
    ch_ref = torch.empty(W, dtype=torch.int64, device=dev)
    sums_ref = torch.empty(4, dtype=torch.float64, device=dev)
    ws = torch.empty(W, dtype=torch.float64, device=dev)
    _lib.call("jqmc_lrdmc_branching", W, _ptr(w), _ptr(vd), _ptr(eL), -12.0, zeta, _ptr(ch_ref), _ptr(sums_ref),
              _ptr(ws), W * 8, stream)
    ch = torch.full((W,), -1, dtype=torch.int64, device=dev)
    sums = torch.full((4,), np.nan, dtype=torch.float64, device=dev)
    _lib.call("jqmc_lrdmc_branching_static", W, _ptr(w), _ptr(vd), _ptr(eL), -12.0, zeta, _ptr(ch), _ptr(sums), stream)
    torch.cuda.synchronize()
    return ch_ref, sums_ref, ch, sums


def test_static_branching_equals_workspace_branching():
    for W, seed in ((50000, 1), (1234, 2)):  # the second call smaller than the first
        ch_ref, sums_ref, ch, sums = _both(W, seed)
        assert torch.equal(ch, ch_ref), W
        assert torch.equal(sums, sums_ref), W
        c = ch.cpu().numpy()
        assert c.min() >= 0 and c.max() < W and np.all(np.diff(c) >= 0)


def test_static_branching_refuses_oversize():
    """Refused on the host before any launch (the arrays are never read)."""
    w = torch.ones(1, dtype=torch.float64, device="cuda")
    ch = torch.empty(1, dtype=torch.int64, device="cuda")
    with pytest.raises(_lib.JqmcError, match="exceeds the static branching scratch"):
        _lib.call("jqmc_lrdmc_branching_static", (1 << 20) + 1, _ptr(w), None, None, 0.0, 0.5, _ptr(ch), None,
                  torch.cuda.current_stream().cuda_stream)
