"""The closure's training-data kernels against closed forms and numpy, and the pipeline end to end
at a small size.

- lbm_box_filter: block means of rho, u and u u against numpy reshape-means (rtol 1e-6).
- lbm_les_samples: on a filtered field built so that the subgrid stress is exactly
  -2 nu0 S + isotropic for a linear velocity field, the target is nu0 and the features are
  Pi = -2 rho cs^2 (tau0 + 3 nu0) S, rho, |u| and the Smagorinsky nu_t of Pi (closed forms).
- lbm_init_fields: uniform fields give lbm_init_equilibrium bit for bit; per-node moments of
  the stored populations reproduce (rho, u).
"""
import ctypes as C

import numpy as np
import pytest

from mojo_lbm_amd import _lib
from tests import _util as U

pytestmark = pytest.mark.gpu


def _st():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_box_filter_block_means(gpu):
    import torch

    rng = np.random.default_rng(0)
    n, f, off = 24, 4, -2
    rho = rng.uniform(0.9, 1.1, (n, n, n)).astype(np.float32)
    u = rng.uniform(-0.1, 0.1, (n, n, n, 3)).astype(np.float32)
    cx = 7  # coarse cell I covers fine 4I-2 .. 4I+1; I = 0 and 6 are partly outside
    out = torch.zeros(cx ** 3 * 10, dtype=torch.float32, device="cuda")
    rd, ud = torch.from_numpy(rho).cuda(), torch.from_numpy(u).cuda()
    _lib.check(_lib.lib().lbm_box_filter(rd.data_ptr(), ud.data_ptr(), 0, n, n, n, f, off, cx, cx, cx,
                                         out.data_ptr(), _st()))
    got = out.cpu().numpy().reshape(cx, cx, cx, 10)
    for I, J, K in [(1, 1, 1), (3, 5, 2), (5, 5, 5), (0, 2, 3), (6, 6, 6)]:
        sl = tuple(slice(max(0, 4 * c + off), min(n, 4 * c + off + 4)) for c in (I, J, K))
        r, v = rho[sl].astype(np.float64), u[sl].astype(np.float64)
        want = [r.mean(), *v.reshape(-1, 3).mean(0)]
        vv = v.reshape(-1, 3)
        want += [np.mean(vv[:, a] * vv[:, b]) for a, b in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]]
        assert np.allclose(got[I, J, K], want, rtol=1e-6, atol=1e-9)


def test_samples_recover_a_known_eddy_viscosity(gpu):
    import torch

    c, nu0, tau0, cs = 12, 2.5e-3, 0.51, 0.1
    G = np.array([[0.0, 2e-3, -1e-3], [1e-3, 5e-4, 0.0], [3e-4, -2e-3, -5e-4]])  # grad u, traceless
    G[2, 2] = -(G[0, 0] + G[1, 1])
    S = 0.5 * (G + G.T)
    idx = np.stack(np.meshgrid(*[np.arange(c)] * 3, indexing="ij"), -1).astype(np.float64)
    ubar = 0.02 + idx @ G.T  # u_d = 0.02 + G[d, a] x_a
    uu = ubar[..., :, None] * ubar[..., None, :] - 2 * nu0 * S + 1e-4 * np.eye(3)
    F = np.zeros((c, c, c, 10))
    F[..., 0] = 1.02
    F[..., 1:4] = ubar
    for k, (a, b) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        F[..., 4 + k] = uu[..., a, b]
    Fd = torch.from_numpy(F.astype(np.float32).ravel()).cuda()
    cells = np.array([(5 * c + 6) * c + 4, (2 * c + 2) * c + 2, 0], np.uint32)
    cd = torch.from_numpy(cells.view(np.int32)).cuda()
    X = torch.zeros(27, dtype=torch.float32, device="cuda")
    y = torch.zeros(3, dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().lbm_les_samples(Fd.data_ptr(), c, c, c, tau0, cs, cd.data_ptr(), 3, X.data_ptr(),
                                          y.data_ptr(), _st()))
    Xh, yh = X.cpu().numpy().reshape(3, 9), y.cpu().numpy()
    assert yh[2] == -1.0  # a border cell is marked
    for i in range(2):
        assert yh[i] == pytest.approx(nu0, rel=2e-3)
        pi = -2 * 1.02 / 3 * (tau0 + 3 * nu0) * np.array([S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]])
        assert np.allclose(Xh[i, :6], pi, rtol=2e-3, atol=1e-9)
        assert Xh[i, 6] == pytest.approx(1.02, rel=1e-6)
        x, yy, z = [(5, 6, 4), (2, 2, 2)][i]
        assert Xh[i, 7] == pytest.approx(np.linalg.norm(ubar[x, yy, z]), rel=1e-5)
        Sb = -np.array(pi) / (2 * 1.02 / 3 * tau0)
        nrm = Sb[0] ** 2 + Sb[3] ** 2 + Sb[5] ** 2 + 2 * (Sb[1] ** 2 + Sb[2] ** 2 + Sb[4] ** 2)
        assert Xh[i, 8] == pytest.approx(cs * cs * np.sqrt(2 * nrm), rel=2e-3)


def test_init_fields(gpu):
    import torch

    desc = U.make_desc("D3Q19", (16, 16, 16), storage=0)
    L = _lib.lib()
    h = _lib.Handle(desc)
    n = 16 ** 3
    rho = torch.full((n,), 1.03, dtype=torch.float32, device="cuda")
    u = torch.tensor([0.01, -0.02, 0.03] * n, dtype=torch.float32, device="cuda")
    fb = L.lbm_f_bytes(C.byref(desc))
    a = torch.empty(fb // 4, dtype=torch.float32, device="cuda")
    b = torch.empty(fb // 4, dtype=torch.float32, device="cuda")
    _lib.check(L.lbm_init_fields(h.ptr, a.data_ptr(), rho.data_ptr(), u.data_ptr(), _st()))
    _lib.check(L.lbm_init_equilibrium(h.ptr, b.data_ptr(), 1.03, _lib.dvec([0.01, -0.02, 0.03]), _st()))
    assert torch.equal(a, b)
    rng = np.random.default_rng(1)
    rh = rng.uniform(0.95, 1.05, n).astype(np.float32)
    uh = rng.uniform(-0.05, 0.05, n * 3).astype(np.float32)
    rd, ud = torch.from_numpy(rh).cuda(), torch.from_numpy(uh).cuda()
    _lib.check(L.lbm_init_fields(h.ptr, a.data_ptr(), rd.data_ptr(), ud.data_ptr(), _st()))
    f = a.cpu().numpy().astype(np.float64)
    amap = U.index_map(desc, 19)  # [x, y, z, q]
    from oracle import oracle as O

    cvec, _, _ = O.lattice(3, 19)
    fq = f[amap]  # [x, y, z, q]
    r = fq.sum(-1)
    mom = fq @ cvec.astype(np.float64)
    assert np.allclose(r.ravel(), rh, rtol=1e-6)
    assert np.allclose((mom / r[..., None]).ravel(), uh, atol=1e-6)


def test_pipeline_end_to_end_small(gpu, tmp_path):
    from mojo_lbm_amd import les

    les.set_sizes(16, re=400.0)  # fine 64^3, coarse 16^3, a Reynolds number these grids resolve
    try:
        X, y, info = les.generate([1], spinup=400, snapshots=2, interval=50, per_snapshot=2000, log=lambda m: None)
        assert len(y) > 100 and np.isfinite(X).all() and (y >= 0).all()
        P, skill = les.train(X, y, steps=60, batch=512, log=lambda m: None)
        assert np.isfinite(P).all() and skill["val_loss_history"]
        res = les.evaluate(P, seed=3, spinup=200, horizon=20, checkpoints=2, log=lambda m: None)
        for row in res["rel_l2"]:
            assert row["smagorinsky"] is not None and row["learned"] is not None
        assert res["coarse_ms_per_step"]["learned"] > 0
    finally:
        les.set_sizes(128)
