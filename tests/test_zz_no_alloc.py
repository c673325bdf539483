"""lbm_mlp_loss_grad_noalloc (static device scratch, no allocation) against lbm_mlp_loss_grad (caller
workspace): bit-identical gradient and loss on two calls in a row, the second with fewer samples, so
the second call runs over the first call's leftover partials."""
import ctypes as C

import numpy as np
import pytest

from mojo_lbm_amd import _lib

pytestmark = pytest.mark.gpu

NPARAM, NTOTAL = 1409, 1446


def _params(seed):
    rng = np.random.default_rng(seed)
    P = np.zeros(NTOTAL, np.float32)
    P[:NPARAM] = (0.3 * rng.standard_normal(NPARAM)).astype(np.float32)
    P[NPARAM:NPARAM + 9] = 0.0  # mean
    P[NPARAM + 9:NPARAM + 18] = 1.0  # inverse std
    P[NPARAM + 18:NPARAM + 27] = -10.0  # training range
    P[NPARAM + 27:NPARAM + 36] = 10.0
    P[NTOTAL - 1] = 1e-3  # nu scale
    return P


def _with_workspace(L, torch, Pd, Xd, yd, n):
    g = torch.full((NPARAM,), float("nan"), dtype=torch.float32, device="cuda")
    loss = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    ws = torch.empty(L.lbm_mlp_workspace_bytes(n), dtype=torch.uint8, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(L.lbm_mlp_loss_grad(Pd.data_ptr(), Xd.data_ptr(), yd.data_ptr(), n, g.data_ptr(), loss.data_ptr(),
                                   ws.data_ptr(), st))
    torch.cuda.synchronize()
    return g.cpu().numpy(), loss.cpu().numpy()


def _no_alloc(L, torch, Pd, Xd, yd, n):
    g = torch.full((NPARAM,), float("nan"), dtype=torch.float32, device="cuda")
    loss = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(L.lbm_mlp_loss_grad_noalloc(Pd.data_ptr(), Xd.data_ptr(), yd.data_ptr(), n, g.data_ptr(),
                                           loss.data_ptr(), st))
    torch.cuda.synchronize()
    return g.cpu().numpy(), loss.cpu().numpy()


def test_noalloc_gradient_matches_workspace_entry_point(gpu):
    import torch

    L = _lib.lib()
    # 100,000 samples fill all 1,024 workgroups (grid-stride); 3,000 use 47 of them
    for call, (n, seed) in enumerate([(100_000, 1), (3_000, 2)]):
        print(f"no-alloc gradient call {call + 1}: n = {n}", flush=True)
        rng = np.random.default_rng(seed)
        X = rng.standard_normal((n, 9)).astype(np.float32)
        y = (np.abs(rng.standard_normal(n)) * 1e-3).astype(np.float32)
        Pd = torch.from_numpy(_params(seed)).cuda()
        Xd, yd = torch.from_numpy(X).cuda(), torch.from_numpy(y).cuda()
        g_ref, l_ref = _with_workspace(L, torch, Pd, Xd, yd, n)
        g, loss = _no_alloc(L, torch, Pd, Xd, yd, n)
        assert np.all(np.isfinite(g_ref)) and np.isfinite(l_ref[0])
        assert np.array_equal(g.view(np.uint32), g_ref.view(np.uint32)), f"gradient differs on call {call + 1}"
        assert np.array_equal(loss.view(np.uint64), l_ref.view(np.uint64)), f"loss differs on call {call + 1}"


def test_noalloc_rejects_bad_arguments(gpu):
    import torch

    L = _lib.lib()
    g = torch.zeros(NPARAM, dtype=torch.float32, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.lbm_mlp_loss_grad_noalloc(None, None, None, 0, g.data_ptr(), None, st) == _lib.ERR_INVALID
    assert "gradient" in L.lbm_last_error().decode()
