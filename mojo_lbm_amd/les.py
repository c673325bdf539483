"""Learned LES closure pipeline for the D3Q19 lid-driven cube at Re = 10,000.

1. Fine runs: a 512^3 Float16C + DDF double-buffer lid-driven cube whose cavity spans
   fine nodes 6..511 (walls at 6 and 511, SOLID padding below 6) so that it is exactly
   4x the 128^3 coarse cavity (walls at 0 and 127): coarse node I covers fine nodes
   4I+3 .. 4I+6. The lid (+Y) moves with u_x = u_lid (1 + 0.1 xi(x, z)), xi a seeded
   smooth field with |xi| <= 1 (the seeded +-10 % lid perturbation).
2. Snapshots: lbm_macroscopic on the fine grid, lbm_box_filter to the coarse grid
   (rho, u, <u_i u_j>), lbm_les_samples for random coarse cells: features and the
   least-squares eddy-viscosity target.
3. Training: lbm_mlp_loss_grad + lbm_mlp_adam on mini-batches (inputs normalised by
   the training mean / std; training range = per-feature min / max).
4. Evaluation: a held-out fine run is filtered at t0; coarse runs with no model,
   Smagorinsky (Cs = 0.1) and the learned closure start from f_eq of that field
   (lbm_init_fields) and are compared with the filtered fine field at later times
   (velocity relative L2 over the coarse fluid nodes); coarse step times are timed
   with HIP events.

Acoustic scaling: u_lid = 0.05 on both grids, so one coarse step is four fine steps and
tau = 0.5 + 3 u_lid N / Re (0.50768 fine, 0.50192 coarse).

Usage: python -m mojo_lbm_amd.les train --out params.json [--spinup K] ...
       python -m mojo_lbm_amd.les evaluate --params params.json --out results.json
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time

import numpy as np

from . import _lib
from .grid import TiledLayouts

U_LID, RE, CS, FACTOR = 0.05, 10000.0, 0.1, 4
N_FINE, N_COARSE, FINE_LO = 512, 128, 6  # fine cavity walls at 6 and 511
OFFSET = FINE_LO - FACTOR + 1  # coarse I covers fine FACTOR*I + OFFSET + (0..3) = 4I+3 .. 4I+6
MLP_NPARAM, MLP_TOTAL = 1409, 1446
MLP_NUSCALE = MLP_TOTAL - 1  # output scale of nu_t in the parameter block


def set_sizes(n_coarse: int, re: float = 10000.0) -> None:
    """Coarse grid size (the fine grid is 4x) and Reynolds number; (128, 10,000) is the
    configuration studied, smaller sizes are for quick checks. tau follows from N and Re."""
    global N_FINE, N_COARSE, RE
    N_COARSE = int(n_coarse)
    N_FINE = FACTOR * N_COARSE
    RE = float(re)


def tau_for(n: int) -> float:
    return 0.5 + 3.0 * U_LID * n / RE


def _torch():
    import torch

    return torch


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def lid_xi(seed: int, X: np.ndarray, Z: np.ndarray) -> np.ndarray:
    """Smooth seeded field on the unit square with max |xi| <= 1: a few low Fourier modes."""
    rng = np.random.default_rng(seed)
    xi = np.zeros(np.broadcast(X, Z).shape)
    for _ in range(6):
        kx, kz = rng.integers(1, 4, size=2)
        xi = xi + rng.uniform(-1, 1) * np.sin(2 * np.pi * (kx * X + rng.uniform())) * np.sin(
            2 * np.pi * (kz * Z + rng.uniform()))
    return xi / 6.0


class Cavity:
    """A lid-driven cube on one handle: cavity nodes lo..n-1 per axis (walls at lo and n-1,
    SOLID below lo), lid at y = n-1 with u_x = u_lid (1 + amp xi), D3Q19 double buffer."""

    def __init__(self, n, lo, tau, storage=0, collision="SRT", les=0, seed=0, amp=0.1, mlp=None, tile=8, cs=None):
        torch = _torch()
        self.n, self.lo, self.tau = n, lo, tau
        ddf = 1 if storage == _lib.STORE_F16C else 0
        coll = {"SRT": _lib.COLL_SRT, "KBC": _lib.COLL_KBC}[collision]
        self.desc = _lib.LbmDesc(D=3, Q=19, nx=n, ny=n, nz=n, tx=tile, ty=tile, tz=tile, method=0, collision=coll,
                                 storage=storage, ddf_shift=ddf, les=les, include_moving_boundary=1,
                                 included_bcs=3, smagorinsky_cs=CS if cs is None else cs)
        self.L = _lib.lib()
        self.h = _lib.Handle(self.desc)
        if mlp is not None:
            p = np.ascontiguousarray(mlp, np.float32)
            _lib.check(self.L.lbm_set_mlp(self.h.ptr, p.ctypes.data, len(p)))
        _lib.check(self.L.lbm_set_option(self.h.ptr, _lib.OPT_STATIC_FLAGS, 1))
        if tile ** 3 > 256:
            # 256-lane z-slab workgroups: at 128^3 every closure runs faster with them (FP32
            # Smagorinsky 0.0715 -> 0.0678 ms per step on MI355X); results are identical
            _lib.check(self.L.lbm_set_option(self.h.ptr, _lib.OPT_STEP_LANES, 256))
        lay = TiledLayouts(3, 19, (n, n, n), (tile, tile, tile))
        ax = np.arange(n)
        solid = ((ax[:, None, None] <= lo) | (ax[None, :, None] <= lo) | (ax[None, None, :] <= lo)
                 | (ax[:, None, None] == n - 1) | (ax[None, :, None] == n - 1) | (ax[None, None, :] == n - 1))
        xs, ys, zs = np.nonzero(solid)
        del solid
        flags = np.zeros(n ** 3, np.uint8)
        flags[lay.index(xs, ys, zs, 0, 1)] = 1
        rdt = np.float64 if storage == _lib.STORE_F64 else np.float32
        bc = np.zeros(n ** 3 * 4, rdt)
        bc.reshape(-1, 4, tile ** 3)[:, 3, :] = 1.0  # wall density rho_w = 1 (the moving-wall term uses it)
        # lid: y = n-1 with lo < x, z < n-1 (edges stay stationary walls)
        lx, lz = np.meshgrid(np.arange(lo + 1, n - 1), np.arange(lo + 1, n - 1), indexing="ij")
        width = n - 2 - lo  # fluid nodes per axis
        X = (lx - lo - 0.5) / width
        Z = (lz - lo - 0.5) / width
        ux = U_LID * (1.0 + amp * lid_xi(seed, X, Z))
        ly = np.full(lx.shape, n - 1)
        bc[lay.index(lx.ravel(), ly.ravel(), lz.ravel(), 0, 4)] = ux.ravel()
        self.flags = torch.from_numpy(flags).cuda()
        self.bc = torch.from_numpy(bc).cuda()
        del flags, bc
        fb = self.L.lbm_f_bytes(C.byref(self.desc))
        self.f = torch.empty(fb, dtype=torch.uint8, device="cuda")
        self.f2 = torch.empty(fb, dtype=torch.uint8, device="cuda")
        _lib.check(self.L.lbm_init_equilibrium(self.h.ptr, self.f.data_ptr(), 1.0, None, _stream()))
        _lib.check(self.L.lbm_init_equilibrium(self.h.ptr, self.f2.data_ptr(), 1.0, None, _stream()))
        cells = n ** 3
        self.rho = torch.empty(cells, dtype=torch.float64 if rdt is np.float64 else torch.float32, device="cuda")
        self.u = torch.empty(cells * 3, dtype=self.rho.dtype, device="cuda")
        self.steps_done = 0

    def step(self, k: int) -> None:
        L, h, st = self.L, self.h.ptr, _stream()
        bcp, flp = self.bc.data_ptr(), self.flags.data_ptr()
        for _ in range(k):
            _lib.check(L.lbm_step_db(h, self.f2.data_ptr(), self.f.data_ptr(), bcp, flp, self.tau, st))
            self.f, self.f2 = self.f2, self.f
        self.steps_done += k

    def macroscopic(self):
        _lib.check(self.L.lbm_macroscopic(self.h.ptr, self.rho.data_ptr(), self.u.data_ptr(), self.f.data_ptr(),
                                          self.bc.data_ptr(), self.flags.data_ptr(), 0, 0, _stream()))
        return self.rho, self.u

    def set_fields(self, rho, u) -> None:
        """f = f_eq(rho, u) per node (both buffers)."""
        for b in (self.f, self.f2):
            _lib.check(self.L.lbm_init_fields(self.h.ptr, b.data_ptr(), rho.data_ptr(), u.data_ptr(), _stream()))


def filtered(fine: Cavity):
    """Coarse (128^3 x 10) box-filtered fields of the fine run's current state (device float32)."""
    torch = _torch()
    rho, u = fine.macroscopic()
    out = torch.empty(N_COARSE ** 3 * 10, dtype=torch.float32, device="cuda")
    _lib.check(fine.L.lbm_box_filter(rho.data_ptr(), u.data_ptr(), fine.desc.storage, fine.n, fine.n, fine.n, FACTOR,
                                     OFFSET, N_COARSE, N_COARSE, N_COARSE, out.data_ptr(), _stream()))
    return out


def samples(coarse, cells_np: np.ndarray):
    torch = _torch()
    cells = torch.from_numpy(np.ascontiguousarray(cells_np, np.uint32).view(np.int32)).cuda()
    n = len(cells_np)
    X = torch.empty(n * 9, dtype=torch.float32, device="cuda")
    y = torch.empty(n, dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().lbm_les_samples(coarse.data_ptr(), N_COARSE, N_COARSE, N_COARSE, tau_for(N_COARSE), CS,
                                          cells.data_ptr(), n, X.data_ptr(), y.data_ptr(), _stream()))
    return X.view(n, 9), y


def _finite(t) -> bool:
    return bool(_torch().isfinite(t).all().item())


def generate(seeds, spinup, snapshots, interval, per_snapshot, collision="SRT", log=print, on_policy=0):
    """Training samples (X [n, 9], y [n]) on the host from fine runs with the given lid seeds.
    on_policy > 0 adds, at every snapshot, samples whose features come from a coarse run
    (Smagorinsky, Cs = 0.1) started from the filtered fine field and advanced alongside it (one
    coarse step per four fine steps), restarted from the filtered field every on_policy
    snapshots while it still tracks the fine run; the targets are the filtered field's. Their
    features span what collide sees in coarse runs, which the filtered field alone does not."""
    rng = np.random.default_rng(12345)
    Xs, ys, info = [], [], []
    for seed in seeds:
        t0 = time.time()
        fine = Cavity(N_FINE, FINE_LO, tau_for(N_FINE), storage=_lib.STORE_F16C, collision=collision, seed=seed)
        chunk = 5000
        done = 0
        while done < spinup:
            k = min(chunk, spinup - done)
            fine.step(k)
            done += k
            c = filtered(fine)
            if not _finite(c):
                raise FloatingPointError(f"fine run seed {seed} diverged by step {fine.steps_done}")
            log(f"# fine seed {seed}: {fine.steps_done} steps, {time.time() - t0:.0f} s")
        coarse = None
        for s in range(snapshots):
            if on_policy and s % on_policy == 0:
                ref = filtered(fine).view(N_COARSE, N_COARSE, N_COARSE, 10)
                coarse = Cavity(N_COARSE, 0, tau_for(N_COARSE), storage=_lib.STORE_F32, les=1, seed=seed)
                coarse.set_fields(ref[..., 0].contiguous().view(-1), ref[..., 1:4].contiguous().view(-1))
            fine.step(interval)
            c = filtered(fine)
            cells = rng.integers(0, N_COARSE ** 3, size=per_snapshot).astype(np.uint32)
            X, y = samples(c, cells)
            Xh, yh = X.cpu().numpy(), y.cpu().numpy()
            keep = (yh >= 0) & np.isfinite(yh) & np.isfinite(Xh).all(axis=1)
            Xs.append(Xh[keep])
            ys.append(yh[keep])
            if coarse is not None:
                coarse.step(interval // FACTOR)
                Xc, fluid = closure_features(coarse)
                ci = _torch().from_numpy(cells.astype(np.int64)).cuda()
                Xo = Xc[ci].cpu().numpy()
                ok = keep & fluid[ci].cpu().numpy() & np.isfinite(Xo).all(axis=1)
                Xs.append(Xo[ok])
                ys.append(yh[ok])
        info.append({"seed": seed, "steps": fine.steps_done, "seconds": round(time.time() - t0, 1),
                     "on_policy_restart": on_policy})
        del coarse
        log(f"# fine seed {seed}: {snapshots} snapshots, {sum(len(a) for a in ys)} samples so far")
        del fine
        _torch().cuda.empty_cache()
    return np.concatenate(Xs), np.concatenate(ys), info


def init_params(X, y, seed=0):
    """Parameter block: fan-in scaled weights, normalisation from the data, range = data min/max."""
    rng = np.random.default_rng(seed)
    P = np.zeros(MLP_TOTAL, np.float64)
    P[0:288] = rng.standard_normal(288) / 3.0
    P[320:1344] = rng.standard_normal(1024) / math.sqrt(32)
    P[1376:1408] = rng.standard_normal(32) / math.sqrt(32)
    mean, std = X.mean(0), X.std(0)
    std[std == 0] = 1.0
    P[1409:1418] = mean
    P[1418:1427] = 1.0 / std
    P[1427:1436] = X.min(0)
    P[1436:1445] = X.max(0)
    scale = float(np.std(y)) if np.std(y) > 0 else 1.0
    P[1445] = scale
    P[1408] = math.log(math.expm1(1.0))  # softplus(b3) = 1: start at the mean target
    return P.astype(np.float32)


def train(X, y, steps=6000, batch=16384, lr=3e-3, seed=0, log=print):
    """Adam on mini-batches with the library's kernels; returns (params, history)."""
    torch = _torch()
    L = _lib.lib()
    rng = np.random.default_rng(seed)
    y = np.minimum(y, np.quantile(y, 0.999))  # the least-squares target has a heavy tail
    perm = rng.permutation(len(y))
    n_val = max(1, len(y) // 10)
    val, tr = perm[:n_val], perm[n_val:]
    P = init_params(X[tr], y[tr], seed)
    Xt = torch.from_numpy(np.ascontiguousarray(X[tr], np.float32)).cuda()
    yt = torch.from_numpy(np.ascontiguousarray(y[tr], np.float32)).cuda()
    Xv = torch.from_numpy(np.ascontiguousarray(X[val], np.float32)).cuda()
    yv = torch.from_numpy(np.ascontiguousarray(y[val], np.float32)).cuda()
    Pd = torch.from_numpy(P).cuda()
    m = torch.zeros(MLP_NPARAM, dtype=torch.float32, device="cuda")
    v = torch.zeros(MLP_NPARAM, dtype=torch.float32, device="cuda")
    g = torch.zeros(MLP_NPARAM, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    ws = torch.empty(max(L.lbm_mlp_workspace_bytes(batch), L.lbm_mlp_workspace_bytes(len(yv))), dtype=torch.uint8,
                     device="cuda")
    n_tr = len(yt)
    nb = max(1, n_tr // batch)
    hist = []
    st = _stream()
    for t in range(1, steps + 1):
        b = int(rng.integers(0, nb))
        lo = b * batch
        n = min(batch, n_tr - lo)
        _lib.check(L.lbm_mlp_loss_grad(Pd.data_ptr(), Xt[lo:].data_ptr(), yt[lo:].data_ptr(), n, g.data_ptr(),
                                       loss.data_ptr(), ws.data_ptr(), st))
        lr_t = lr * (0.1 + 0.9 * 0.5 * (1.0 + math.cos(math.pi * t / steps)))
        _lib.check(L.lbm_mlp_adam(Pd.data_ptr(), m.data_ptr(), v.data_ptr(), g.data_ptr(), t, lr_t, 0.9, 0.999,
                                  1e-8, st))
        if t % 500 == 0 or t == steps:
            _lib.check(L.lbm_mlp_loss_grad(Pd.data_ptr(), Xv.data_ptr(), yv.data_ptr(), len(yv), g.data_ptr(),
                                           loss.data_ptr(), ws.data_ptr(), st))
            vl = float(loss.item())
            hist.append((t, vl))
            log(f"# train step {t}: validation loss {vl:.4e}")
    P = Pd.cpu().numpy()
    # a-priori skill on the validation set: correlation with the target, closure vs Smagorinsky
    nu = torch.empty(len(yv), dtype=torch.float32, device="cuda")
    _lib.check(L.lbm_mlp_forward(Pd.data_ptr(), Xv.data_ptr(), len(yv), nu.data_ptr(), st))
    nu_h, yv_h = nu.cpu().numpy(), y[val]
    smag = X[val][:, 8]
    skill = {
        "corr_closure": float(np.corrcoef(nu_h, yv_h)[0, 1]),
        "corr_smagorinsky": float(np.corrcoef(smag, yv_h)[0, 1]),
        "rel_l2_closure": float(np.linalg.norm(nu_h - yv_h) / np.linalg.norm(yv_h)),
        "rel_l2_smagorinsky": float(np.linalg.norm(smag - yv_h) / np.linalg.norm(yv_h)),
        "val_loss_history": hist,
        "n_train": int(n_tr), "n_val": int(len(yv)),
    }
    return P, skill


def closure_features(c: Cavity):
    """The nine features collide computes for the learned closure at every node of a coarse
    cavity's current state (FP32 storage): the populations are pulled from x - c_q (halfway
    bounce-back from SOLID neighbours plus the moving-wall term 6 w rho_w (c . u_w)), then
    Pi_neq (xx, xy, xz, yy, yz, zz), rho, |u| and the Smagorinsky nu_t (Cs = 0.1) of that
    Pi_neq. Returns (features (n^3, 9) float32 in row-major node order, fluid mask (n^3,)),
    both device tensors; rows of SOLID nodes are not meaningful."""
    torch = _torch()
    from .lattice import get_D3Q19

    lat = get_D3Q19()
    n, tile = c.n, c.desc.tx
    lay = TiledLayouts(3, 19, (n, n, n), (tile, tile, tile))
    idx = torch.from_numpy(lay.index_map(19)).cuda()
    f = c.f.view(torch.float32)[idx]  # (n, n, n, 19) post-collision populations
    bidx = torch.from_numpy(lay.index_map(4)).cuda()
    bc = c.bc.view(torch.float32)[bidx]
    solid = (c.flags[torch.from_numpy(lay.index_map(1)[..., 0]).cuda()] & 0x0F) == 1
    cq = lat.directions
    w = torch.tensor(lat.weights, dtype=torch.float32, device="cuda")
    opp = lat.opposite_indices
    fin = torch.empty_like(f)
    for q in range(19):
        sh = tuple(int(v) for v in cq[q])
        pulled = torch.roll(f[..., q], shifts=sh, dims=(0, 1, 2))  # value at x - c_q
        nb_solid = torch.roll(solid, shifts=sh, dims=(0, 1, 2))
        nb_bc = torch.roll(bc, shifts=sh, dims=(0, 1, 2))
        mw = 6.0 * w[q] * nb_bc[..., 3] * (nb_bc[..., 0] * sh[0] + nb_bc[..., 1] * sh[1] + nb_bc[..., 2] * sh[2])
        fin[..., q] = torch.where(nb_solid, f[..., int(opp[q])] + (mw if q else 0.0), pulled)
    fl = fin.reshape(-1, 19)
    cqt = torch.tensor(cq, dtype=torch.float32, device="cuda")
    rho = fl.sum(1)
    u = fl @ cqt / rho[:, None]
    uu = (u * u).sum(1)
    eu = u @ cqt.T
    feq = w[None, :] * rho[:, None] * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uu[:, None])
    fn = fl - feq
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    pi = torch.stack([fn @ (cqt[:, a] * cqt[:, b]) for a, b in pairs], 1)
    k = -1.0 / (2 * rho * (1.0 / 3.0) * c.tau)
    S = pi * k[:, None]
    snorm = (S * S * torch.tensor([1, 2, 2, 1, 2, 1], dtype=torch.float32, device="cuda")).sum(1)
    nu_s = CS * CS * torch.sqrt(2 * snorm)
    X = torch.cat([pi, rho[:, None], torch.sqrt(uu)[:, None], nu_s[:, None]], 1).contiguous()
    return X, ~solid.reshape(-1)


def fallback_fraction(c: Cavity, params) -> dict:
    """Share of fluid nodes whose features lie outside the closure's training range (where
    collide uses the Smagorinsky value), per feature and in total, and of 64-node waves in
    which every node falls back."""
    torch = _torch()
    X, fluid = closure_features(c)
    X = X[fluid]
    P = torch.from_numpy(np.asarray(params, np.float32)).cuda()
    lo, hi = P[1427:1436], P[1436:1445]
    out = (X < lo) | (X > hi)
    names = ["pi_xx", "pi_xy", "pi_xz", "pi_yy", "pi_yz", "pi_zz", "rho", "u", "nu_smag"]
    return {"nodes": int(X.shape[0]), "fallback": float(out.any(1).float().mean()),
            "per_feature": {nm: float(out[:, i].float().mean()) for i, nm in enumerate(names)}}


def _coarse_u(c: Cavity):
    rho, u = c.macroscopic()
    return u.view(N_COARSE, N_COARSE, N_COARSE, 3)[1:-1, 1:-1, 1:-1].double()


def scaled_params(params, factor):
    """The closure with its nu_t output scaled by factor (an a-posteriori calibration)."""
    p = np.array(params, np.float32, copy=True)
    p[MLP_NUSCALE] = np.float32(p[MLP_NUSCALE] * factor)
    return p


def evaluate(params, seed, spinup, horizon, checkpoints=4, collision="SRT", log=print, stats_steps=0,
             sample_every=25, nu_scales=(), cs_values=()):
    """A-posteriori comparison of coarse closures against the filtered held-out fine run: velocity
    rel-L2 at checkpoints, then (stats_steps > 0) the rel-L2 of the time-mean velocity and of the
    rms velocity fluctuation over stats_steps further coarse steps sampled every sample_every steps
    (instantaneous fields decorrelate from the fine run within about a thousand coarse steps; the
    statistics do not). Each factor of nu_scales adds a coarse run of the closure with its nu_t
    scaled by that factor ("learned_x<factor>"), each value of cs_values a Smagorinsky run with
    that constant ("smagorinsky_cs<value>"). At every checkpoint the share of the learned run's
    nodes that fall back to Smagorinsky is recorded."""
    torch = _torch()
    t0 = time.time()
    fine = Cavity(N_FINE, FINE_LO, tau_for(N_FINE), storage=_lib.STORE_F16C, collision=collision, seed=seed)
    done = 0
    while done < spinup:
        k = min(5000, spinup - done)
        fine.step(k)
        done += k
        log(f"# eval fine: {fine.steps_done} steps, {time.time() - t0:.0f} s")
    ref0 = filtered(fine).view(N_COARSE, N_COARSE, N_COARSE, 10)
    if not _finite(ref0):
        raise FloatingPointError("evaluation fine run diverged")
    rho0 = ref0[..., 0].contiguous().view(-1)
    u0 = ref0[..., 1:4].contiguous().view(-1)
    models = {"none": (0, None), "smagorinsky": (1, None), "learned": (2, params)}
    for fct in nu_scales:
        models[f"learned_x{fct:g}"] = (2, scaled_params(params, fct))
    for cs in cs_values:
        models[f"smagorinsky_cs{cs:g}"] = (1, None, cs)
    coarse = {}
    for name, spec in models.items():
        les, p = spec[0], spec[1]
        c = Cavity(N_COARSE, 0, tau_for(N_COARSE), storage=_lib.STORE_F32, les=les, seed=seed, mlp=p,
                   cs=spec[2] if len(spec) > 2 else CS)
        c.set_fields(rho0, u0)
        coarse[name] = c
    per = horizon // checkpoints
    rows = []
    ref_sum = None
    sums = {name: None for name in coarse}
    for k in range(1, checkpoints + 1):
        fine.step(FACTOR * per)
        ref = filtered(fine).view(N_COARSE, N_COARSE, N_COARSE, 10)[1:-1, 1:-1, 1:-1, 1:4].double()
        row = {"coarse_steps": k * per, "reference_finite": bool(torch.isfinite(ref).all().item())}
        ref_sum = ref.clone() if ref_sum is None else ref_sum + ref
        for name, c in coarse.items():
            c.step(per)
            u = _coarse_u(c)
            sums[name] = u.clone() if sums[name] is None else sums[name] + u
            err = float(torch.linalg.vector_norm(u - ref) / torch.linalg.vector_norm(ref))
            row[name] = err if math.isfinite(err) else None
        row["learned_fallback"] = fallback_fraction(coarse["learned"], params)
        rows.append(row)
        log(f"# eval checkpoint {row}")
    # time-averaged velocity over the checkpoints
    mean_err = {}
    for name, acc in sums.items():
        e = float(torch.linalg.vector_norm(acc - ref_sum) / torch.linalg.vector_norm(ref_sum))
        mean_err[name] = e if math.isfinite(e) else None
    stats = None
    if stats_steps > 0:
        n = 0
        acc = {name: [0.0, 0.0] for name in ["reference", *coarse]}

        def add(name, u):
            a = acc[name]
            a[0] = a[0] + u
            a[1] = a[1] + u * u

        for k in range(stats_steps // sample_every):
            fine.step(FACTOR * sample_every)
            add("reference", filtered(fine).view(N_COARSE, N_COARSE, N_COARSE, 10)[1:-1, 1:-1, 1:-1, 1:4].double())
            for name, c in coarse.items():
                c.step(sample_every)
                add(name, _coarse_u(c))
            n += 1
            if (k + 1) % 40 == 0:
                log(f"# eval statistics: {n} samples, {time.time() - t0:.0f} s")
        mean_r = acc["reference"][0] / n
        rms_r = (acc["reference"][1] / n - mean_r * mean_r).clamp_min(0).sqrt()
        stats = {"coarse_steps": stats_steps, "samples": n, "rel_l2_mean_velocity": {}, "rel_l2_rms_fluctuation": {}}
        for name in coarse:
            mean_c = acc[name][0] / n
            rms_c = (acc[name][1] / n - mean_c * mean_c).clamp_min(0).sqrt()
            e1 = float(torch.linalg.vector_norm(mean_c - mean_r) / torch.linalg.vector_norm(mean_r))
            e2 = float(torch.linalg.vector_norm(rms_c - rms_r) / torch.linalg.vector_norm(rms_r))
            stats["rel_l2_mean_velocity"][name] = e1 if math.isfinite(e1) else None
            stats["rel_l2_rms_fluctuation"][name] = e2 if math.isfinite(e2) else None
        log(f"# eval statistics {stats}")
    times = {}
    for name in ("smagorinsky", "learned", "none"):
        c = coarse[name]
        c.step(5)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        c.step(200)
        end.record()
        torch.cuda.synchronize()
        times[name] = start.elapsed_time(end) / 200
    return {"seed": seed, "fine_steps_spinup": spinup, "rel_l2": rows, "rel_l2_time_mean": mean_err, "statistics": stats,
            "coarse_ms_per_step": times,
            "learned_over_smagorinsky_step_time": times["learned"] / times["smagorinsky"],
            "seconds": round(time.time() - t0, 1)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--seeds", default="1,2")
    t.add_argument("--spinup", type=int, default=30000)
    t.add_argument("--snapshots", type=int, default=20)
    t.add_argument("--interval", type=int, default=500)
    t.add_argument("--per-snapshot", type=int, default=60000)
    t.add_argument("--steps", type=int, default=6000)
    t.add_argument("--collision", default="SRT")
    t.add_argument("--coarse", type=int, default=128)
    t.add_argument("--on-policy", type=int, default=0,
                   help="add coarse-run features, restarting the coarse run every N snapshots (0: off)")
    t.add_argument("--out", required=True)
    e = sub.add_parser("evaluate")
    e.add_argument("--params", required=True)
    e.add_argument("--seed", type=int, default=3)
    e.add_argument("--spinup", type=int, default=30000)
    e.add_argument("--horizon", type=int, default=2000, help="coarse steps compared")
    e.add_argument("--collision", default="SRT")
    e.add_argument("--coarse", type=int, default=128)
    e.add_argument("--stats-steps", type=int, default=0, help="coarse steps of time statistics after the checkpoints")
    e.add_argument("--sample-every", type=int, default=25)
    e.add_argument("--nu-scales", default="", help="comma-separated factors: extra closure runs with nu_t scaled")
    e.add_argument("--cs-values", default="", help="comma-separated Smagorinsky constants: extra Smagorinsky runs")
    e.add_argument("--out", required=True)
    sc = sub.add_parser("scale", help="write a copy of a parameter file with the nu_t output scaled")
    sc.add_argument("--params", required=True)
    sc.add_argument("--factor", type=float, required=True)
    sc.add_argument("--note", default="")
    sc.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    if a.cmd != "scale":
        set_sizes(a.coarse)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if a.cmd == "train":
        seeds = [int(s) for s in a.seeds.split(",")]
        try:
            X, y, info = generate(seeds, a.spinup, a.snapshots, a.interval, a.per_snapshot, a.collision, log,
                                  a.on_policy)
        except FloatingPointError as exc:
            if a.collision != "SRT":
                raise
            log(f"# {exc}; the fine runs are repeated with KBC collision")
            a.collision = "KBC"
            X, y, info = generate(seeds, a.spinup, a.snapshots, a.interval, a.per_snapshot, a.collision, log,
                                  a.on_policy)
        log(f"# {len(y)} samples; target mean {y.mean():.3e}, max {y.max():.3e}")
        P, skill = train(X, y, a.steps, log=log)
        out = {"params": [float(v) for v in P], "skill": skill, "runs": info,
               "config": {"fine": N_FINE, "coarse": N_COARSE, "u_lid": U_LID, "re": RE, "cs": CS,
                          "tau_fine": tau_for(N_FINE), "tau_coarse": tau_for(N_COARSE), "spinup": a.spinup,
                          "snapshots": a.snapshots, "interval": a.interval, "per_snapshot": a.per_snapshot,
                          "collision": a.collision, "on_policy": a.on_policy}}
        with open(a.out, "w") as fh:
            json.dump(out, fh)
        log(json.dumps({k: v for k, v in skill.items() if k != "val_loss_history"}))
    elif a.cmd == "scale":
        saved = json.load(open(a.params))
        saved["params"] = [float(v) for v in scaled_params(saved["params"], a.factor)]
        saved.setdefault("calibration", []).append({"nu_scale_factor": a.factor, "note": a.note})
        with open(a.out, "w") as fh:
            json.dump(saved, fh)
    else:
        saved = json.load(open(a.params))
        P = np.asarray(saved["params"], np.float32)
        coll = saved.get("config", {}).get("collision", a.collision)
        scales = [float(v) for v in a.nu_scales.split(",") if v.strip()]
        css = [float(v) for v in a.cs_values.split(",") if v.strip()]
        res = evaluate(P, a.seed, a.spinup, a.horizon, collision=coll, log=log, stats_steps=a.stats_steps,
                       sample_every=a.sample_every, nu_scales=scales, cs_values=css)
        res["fine_collision"] = coll
        with open(a.out, "w") as fh:
            json.dump(res, fh)
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
