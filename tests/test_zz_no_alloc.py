"""The allocation-free gradient entry point (dfsph_mlp_grad_noalloc, for callers that cannot allocate device memory)
against dfsph_mlp_grad with a caller workspace, through the C-ABI: the same kernels and the same fixed-order
reduction, so the two agree bit for bit; two calls in a row, the second with fewer samples, share the library's
static partials array."""

from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

MAX_SAMPLES = 4096  # DFSPH_MLP_NOALLOC_MAX_SAMPLES


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def test_noalloc_gradient_matches_the_workspace_gradient_bitwise():
    import torch

    from alloy_amd.sph import _abi
    from alloy_amd.sph.learned import WarmStartMLP

    lib = _abi.load_library()
    net = WarmStartMLP("cuda", seed=5)
    P = int(lib.dfsph_mlp_param_count())
    gen = torch.Generator(device="cuda").manual_seed(11)
    for n in (3000, 700):  # the second call reads fewer partials than the first left behind
        print(f"[no_alloc] n={n}", flush=True)
        sys.stdout.flush()
        X = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
        X[:, :13] = torch.randn((n, 13), generator=gen, device="cuda")
        T = torch.randn((n, 2), generator=gen, device="cuda")
        g_ws = torch.full((P + 1,), float("nan"), device="cuda")
        g_na = torch.full((P + 1,), float("nan"), device="cuda")
        ws = torch.empty(int(lib.dfsph_mlp_grad_scratch_floats(n)), dtype=torch.float32, device="cuda")
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert lib.dfsph_mlp_grad(_vp(net.params), _vp(X), _vp(T), n, n, _vp(ws), ws.numel(), _vp(g_ws), st) == 0
        assert lib.dfsph_mlp_grad_noalloc(_vp(net.params), _vp(X), _vp(T), n, n, _vp(g_na), st) == 0
        torch.cuda.synchronize()
        assert torch.isfinite(g_ws).all() and torch.count_nonzero(g_ws[:P]) > 0
        assert torch.equal(g_ws, g_na), float((g_ws - g_na).abs().max())


def test_noalloc_gradient_refuses_more_than_its_capacity():
    import torch

    from alloy_amd.sph import _abi

    lib = _abi.load_library()
    P = int(lib.dfsph_mlp_param_count())
    n = MAX_SAMPLES + 1
    X = torch.zeros((n, 16), device="cuda")
    T = torch.zeros((n, 2), device="cuda")
    g = torch.zeros(P + 1, device="cuda")
    p = torch.zeros(P, device="cuda")
    assert lib.dfsph_mlp_grad_noalloc(_vp(p), _vp(X), _vp(T), n, n, _vp(g), None) == _abi.DFSPH_ECAPACITY
    assert lib.dfsph_mlp_grad_noalloc(_vp(p), _vp(X), _vp(T), 0, 1, _vp(g), None) == _abi.DFSPH_EINVAL
