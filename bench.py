"""Stream-collide throughput on one MI355X: D3Q19 256^3 lid-driven cube, MLUPs and
fraction of the HBM roofline.

Headline line (value): D3Q19 SRT, 256^3, FP32 populations, two-buffer pull
streaming on the (x,y,z,q) layout with 8x8x8 tiles, bounce-back walls plus a
moving lid u = 0.1 on +Y, Re = 100 (tau = 0.5 + 3 u N / Re), started from
equilibrium at rho = 1, u = 0. K timed steps after W warm-up steps; inputs
resident in HBM; timed with HIP events on the stream the kernels run on.

The same line carries the Float16C + DDF-shift (double buffer and single-buffer
esoteric pull) and FP32 esoteric-pull variants of the same case under "variants",
the roofline of the dominant kernel, and the oracle (C++ restatement of the
reference's per-node step, OpenMP over the host cores) timed on a bounded
sample of the same workload as the CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--only fp32]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def algorithmic_bytes(Q, nx, ny, nz, s, walls=True):
    """Bytes per step: read + write every population of every fluid node, plus one flag byte per node."""
    n_c = nx * ny * nz
    n_f = (nx - 2) * (ny - 2) * (nz - 2) if walls else n_c
    return n_f * 2 * Q * s + n_c


def make_run(desc, u_lid, walls=True):
    import torch

    from mojo_lbm_amd import _lib

    L = _lib.lib()
    h = _lib.Handle(desc)
    fb, bb, gb = L.lbm_f_bytes(C.byref(desc)), L.lbm_bc_bytes(C.byref(desc)), L.lbm_flags_bytes(C.byref(desc))
    f = torch.empty(fb, dtype=torch.uint8, device="cuda")
    f2 = torch.empty(fb, dtype=torch.uint8, device="cuda")
    bc = torch.zeros(bb, dtype=torch.uint8, device="cuda")
    flags = torch.zeros(gb, dtype=torch.uint8, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(L.lbm_init_equilibrium(h.ptr, f.data_ptr(), 1.0, None, st))
    _lib.check(L.lbm_init_equilibrium(h.ptr, f2.data_ptr(), 1.0, None, st))
    D = desc.D
    lid = _lib.dvec([u_lid] + [0.0] * (D - 1))
    zero = _lib.dvec([0.0] * D)
    if walls:
        _lib.check(L.lbm_set_exterior_wall(h.ptr, flags.data_ptr(), bc.data_ptr(), b"+Y", 1, lid, 1.0, st))
        for side in (b"-Y", b"+X", b"-X", b"-Z", b"+Z")[: (3 if D == 2 else 5)]:
            _lib.check(L.lbm_set_exterior_wall(h.ptr, flags.data_ptr(), bc.data_ptr(), side, 1, zero, 1.0, st))
    torch.cuda.synchronize()
    return dict(h=h, f=f, f2=f2, bc=bc, flags=flags, L=L, st=st, desc=desc)


def time_run(run, tau, steps, warmup):
    """Runs warmup + steps launches; returns seconds for the timed steps (HIP events on the launch stream)."""
    import torch

    from mojo_lbm_amd import _lib

    L, h, st, d = run["L"], run["h"], run["st"], run["desc"]
    bufs = [run["f"], run["f2"]]
    bc, flags = run["bc"].data_ptr(), run["flags"].data_ptr()
    state = {"i": 0}

    def one(k):
        if d.method == 0:
            a, b = bufs[k % 2], bufs[(k + 1) % 2]
            return L.lbm_step_db(h.ptr, b.data_ptr(), a.data_ptr(), bc, flags, tau, st)
        return L.lbm_step_ep(h.ptr, bufs[0].data_ptr(), bc, flags, tau, int(k % 2 == 0), st)

    for k in range(warmup):
        _lib.check(one(k))
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    k0 = warmup + (warmup % 2)  # keep the double-buffer parity aligned
    for k in range(k0, k0 + steps):
        rc = one(k)
        if rc:
            _lib.check(rc)
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1e3


def cpu_baseline(desc, tau, budget_s=12.0):
    """Oracle (C++ restatement of the reference's per-node step, OpenMP) on the host cores:
    256^3 lid-driven cube, as many full steps as fit in the budget (at least one)."""
    import numpy as np

    from oracle import oracle as O

    threads = O.max_threads()
    O.set_threads(threads)
    f, bc, flags = O.alloc(desc)
    O.init_equilibrium(desc, f)
    D = desc.D
    O.set_exterior_wall(desc, flags, bc, "+Y", 1, [0.1] + [0.0] * (D - 1), 1.0)
    for s in ["-Y", "+X", "-X", "-Z", "+Z"][: (3 if D == 2 else 5)]:
        O.set_exterior_wall(desc, flags, bc, s, 1, [0.0] * D, 1.0)
    f2 = f.copy()
    t0 = time.perf_counter()
    O.step_db(desc, f2, f, bc, flags, tau)
    t1 = time.perf_counter()
    per = t1 - t0
    n_more = max(0, min(20, int(budget_s / per) - 1))
    t0 = time.perf_counter()
    for k in range(n_more):
        if k % 2 == 0:
            O.step_db(desc, f, f2, bc, flags, tau)
        else:
            O.step_db(desc, f2, f, bc, flags, tau)
    t1 = time.perf_counter()
    if n_more:
        per = (t1 - t0) / n_more
    # the same restatement on one thread: one step
    O.set_threads(1)
    t0 = time.perf_counter()
    O.step_db(desc, f, f2, bc, flags, tau)
    per1 = time.perf_counter() - t0
    O.set_threads(threads)
    n_c = desc.nx * desc.ny * desc.nz
    del f, f2, bc, flags
    return {
        "value": round(n_c / per / 1e6, 2),
        "unit": "MLUPs",
        "cores": threads,
        "kind": "port",
        "sample": f"{desc.nx}^3 D3Q19 FP32 lid-driven cube, {1 + n_more} oracle steps (1 untimed), "
                  f"OpenMP over {threads} threads",
        "single_thread": {"value": round(n_c / per1 / 1e6, 3), "unit": "MLUPs", "cores": 1,
                          "sample": f"{desc.nx}^3, one oracle step on one thread"},
    }


def accuracy_study(N, checkpoints, u_lid=0.1, Re=1000.0):
    """Float16C + DDF (double buffer and esoteric pull) and FP32 runs of the D3Q19 lid-driven cube
    against FP64 runs of the same case and method: velocity relative L2 and max error at each
    checkpoint, and the relative drift of the FP64-reduced total mass."""
    import numpy as np
    import torch

    from mojo_lbm_amd import _lib
    from mojo_lbm_amd._lib import LbmDesc

    tau = 0.5 + 3.0 * u_lid * N / Re
    L = _lib.lib()

    def desc_for(storage, method, ddf):
        return LbmDesc(D=3, Q=19, nx=N, ny=N, nz=N, tx=8, ty=8, tz=8, method=method, collision=0,
                       storage=storage, ddf_shift=ddf, les=0, include_moving_boundary=1, included_bcs=3,
                       smagorinsky_cs=0.1)

    def trajectory(desc):
        run = make_run(desc, u_lid, True)
        st, h = run["st"], run["h"]
        dt = torch.float64 if desc.storage == 2 else torch.float32
        rho = torch.empty(N ** 3, dtype=dt, device="cuda")
        u = torch.empty(N ** 3 * 3, dtype=dt, device="cuda")
        tot = torch.zeros(4, dtype=torch.float64, device="cuda")
        bufs = [run["f"], run["f2"]]
        out, k = {}, 0

        def cur():
            return bufs[k % 2] if desc.method == 0 else bufs[0]

        def moments():
            _lib.check(L.lbm_macroscopic(h.ptr, rho.data_ptr(), u.data_ptr(), cur().data_ptr(), run["bc"].data_ptr(),
                                         run["flags"].data_ptr(), int(k % 2 == 0), 0, st))
            _lib.check(L.lbm_reduce_moments_f64(h.ptr, tot.data_ptr(), cur().data_ptr(), run["bc"].data_ptr(),
                                                run["flags"].data_ptr(), int(k % 2 == 0), st))
            torch.cuda.synchronize()
            return u.double().cpu().numpy().reshape(N, N, N, 3), float(tot[0].item())

        _, m0 = moments()
        for target in checkpoints:
            while k < target:
                if desc.method == 0:
                    _lib.check(L.lbm_step_db(h.ptr, bufs[(k + 1) % 2].data_ptr(), bufs[k % 2].data_ptr(),
                                             run["bc"].data_ptr(), run["flags"].data_ptr(), tau, st))
                else:
                    _lib.check(L.lbm_step_ep(h.ptr, bufs[0].data_ptr(), run["bc"].data_ptr(), run["flags"].data_ptr(),
                                             tau, int(k % 2 == 0), st))
                k += 1
            uu, m = moments()
            out[target] = (uu, (m - m0) / m0)
            print(f"# accuracy: storage={desc.storage} method={desc.method} step {target} mass drift {(m - m0) / m0:.3e}",
                  file=sys.stderr, flush=True)
        del run
        torch.cuda.empty_cache()
        return out

    fluid = np.zeros((N, N, N), bool)
    fluid[1:-1, 1:-1, 1:-1] = True
    results = {}
    for method, mname in ((0, "double_buffer"), (1, "esoteric_pull")):
        ref = trajectory(desc_for(2, method, 0))
        cases = [("f16c_ddf", desc_for(1, method, 1))] + ([("fp32", desc_for(0, method, 0))] if method == 0 else [])
        for cname, d in cases:
            tr = trajectory(d)
            rows = {}
            for t in checkpoints:
                u64, drift64 = ref[t]
                u, drift = tr[t]
                e = (u - u64)[fluid]
                rows[str(t)] = {
                    "u_rel_l2": float(np.linalg.norm(e) / np.linalg.norm(u64[fluid])),
                    "u_max_abs": float(np.max(np.abs(e))),
                    "u_max_ref": float(np.max(np.abs(u64[fluid]))),
                    "mass_drift": drift,
                    "mass_drift_fp64_run": drift64,
                }
            results[f"{cname}_{mname}"] = rows
            del tr
        del ref
    line = {"study": f"D3Q19 lid-driven cube {N}^3, u_lid={u_lid}, Re={Re:g}, tau={tau:.4f}: error vs the FP64 run "
                     "of the same case and method (fluid nodes)", "results": results}
    return line


def cylinder_batch(steps, autotune_steps=100, seed=2024, compare=False, tile=None):
    """64 seeded 2-D cylinder channels (D2Q9 SRT FP32, interpolated bounce-back) advanced together:
    one launch per step for all cases, hipGraph replays. Times `steps` steps from rest after the
    autotuner has picked the tile shape and grid; returns aggregate MLUPs over the real (unpadded)
    nodes and the per-case drag / lift coefficients. compare=True also runs every case alone
    through lbm_step_db and compares its populations and coefficients."""
    import numpy as np
    import torch

    from mojo_lbm_amd import DoubleBufferConfig, Flags, get_D2Q9
    from mojo_lbm_amd.batch import CaseBatch
    from mojo_lbm_amd.cylinder import random_cases

    lat = get_D2Q9()
    cases = random_cases(64, seed)
    b = CaseBatch(lat, [(c.nx, c.ny, 1) for c in cases], [c.tau for c in cases],
                  DoubleBufferConfig(BCs={Flags.EQUILIBRIUM}))
    for i, c in enumerate(cases):
        flags, bc, links = c.arrays(lat)
        b.set_case(i, flags, bc, links)
    t0 = time.time()
    if tile is None:
        tile, grid = b.autotune(autotune_steps)
    else:  # a fixed tile shape (profiling runs that should hold one kernel configuration only)
        b.set_tile(tile)
        grid = 0
    tune_s = time.time() - t0
    info = b.info()
    b.init_equilibrium()
    b.run(info["graph_steps"])  # captures the graph outside the timed region
    b.init_equilibrium()
    f0 = [b.get_f(i).cpu().numpy() for i in range(len(cases))] if compare else None
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    b.run(steps)
    end.record()
    torch.cuda.synchronize()
    t = start.elapsed_time(end) / 1e3
    forces = b.force()
    coef = np.array([c.coefficients(f) for c, f in zip(cases, forces)])
    real = info["real_nodes"]
    bytes_step = real * 2 * lat.Q * 4 + info["padded_nodes"]
    out = {
        "workload": f"64 D2Q9 SRT FP32 cylinder channels (seed {seed}), nx 200..800, interpolated bounce-back, "
                    f"{steps} steps each from rest, one launch per step for all cases",
        "cases": len(cases), "real_nodes": real, "padded_nodes": info["padded_nodes"], "steps": steps,
        "tile": list(tile), "grid_blocks": grid, "graph_steps": info["graph_steps"],
        "autotune_s": round(tune_s, 2),
        "seconds": round(t, 4), "ms_per_step": round(t / steps * 1e3, 5),
        "mlups": round(real * steps / t / 1e6, 1),
        "roofline_frac": round(bytes_step / (t / steps) / 1e9 / HBM_PEAK_GBS, 4),
        "tau_range": [round(min(c.tau for c in cases), 5), round(max(c.tau for c in cases), 5)],
        "unstable_cases": int(np.sum(~np.isfinite(coef).all(axis=1))),
        "cd": [float(np.nanmin(coef[:, 0])), float(np.nanmedian(coef[:, 0])), float(np.nanmax(coef[:, 0]))],
        "cl_abs_max": float(np.nanmax(np.abs(coef[:, 1]))),
    }
    if compare:
        from mojo_lbm_amd import _lib

        L = _lib.lib()
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        same, dcd, dcl, t_single = 0, 0.0, 0.0, 0.0
        per_case = []
        for i, c in enumerate(cases):
            d, flags, bc, nodes, qd = b.export_case(i)
            h = _lib.Handle(d)
            _lib.check(L.lbm_set_option(h.ptr, _lib.OPT_STATIC_FLAGS, 1))
            _lib.check(L.lbm_set_links(h.ptr, nodes.ctypes.data, qd.ctypes.data, len(nodes)))
            fa = torch.from_numpy(f0[i]).cuda()
            fb = fa.clone()
            bct = torch.from_numpy(bc).cuda()
            flt = torch.from_numpy(flags).cuda()
            torch.cuda.synchronize()
            t1 = time.time()
            for _ in range(steps):
                _lib.check(L.lbm_step_db(h.ptr, fb.data_ptr(), fa.data_ptr(), bct.data_ptr(), flt.data_ptr(),
                                         c.tau, st))
                fa, fb = fb, fa
            torch.cuda.synchronize()
            t_single += time.time() - t1
            fo = torch.zeros(3, dtype=torch.float64, device="cuda")
            _lib.check(L.lbm_link_force(h.ptr, fo.data_ptr(), st))
            cs = c.coefficients(fo.cpu().numpy())
            eq = bool(torch.equal(fa.view(torch.uint8), b.get_f(i).view(torch.uint8)))
            same += eq
            dcd = max(dcd, abs(cs[0] - coef[i, 0]))
            dcl = max(dcl, abs(cs[1] - coef[i, 1]))
            per_case.append({"nx": c.nx, "ny": c.ny, "D": round(c.diameter, 3), "Re": round(c.re, 2),
                             "tau": round(c.tau, 5), "cd_batch": coef[i, 0], "cl_batch": coef[i, 1],
                             "cd_single": cs[0], "cl_single": cs[1], "bitwise_equal": eq})
            del h
        out["compare"] = {"bitwise_equal_cases": same, "max_abs_dcd": dcd, "max_abs_dcl": dcl,
                          "single_runs_seconds": round(t_single, 2), "per_case": per_case}
    b.close()
    return out


def les_step_times(steps=500):
    """128^3 D3Q19 lid-driven cube (Re = 10,000, FP32): step time with the Smagorinsky closure and with
    the learned closure (the trained parameters shipped in mojo_lbm_amd/data), HIP events."""
    import numpy as np
    import torch

    from mojo_lbm_amd import _lib, les

    params = np.asarray(json.loads((ROOT / "mojo_lbm_amd" / "data" / "les_closure_d3q19_re10000.json")
                                   .read_text())["params"], np.float32)
    out = {}
    for name, les_id, p in (("none", 0, None), ("smagorinsky", 1, None), ("learned", 2, params)):
        c = les.Cavity(128, 0, les.tau_for(128), storage=_lib.STORE_F32, les=les_id, seed=3, mlp=p)
        c.step(20)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        c.step(steps)
        end.record()
        torch.cuda.synchronize()
        out[name] = round(start.elapsed_time(end) / steps, 5)
        del c
    return {"workload": "D3Q19 lid-driven cube 128^3 FP32, Re 10,000, double buffer", "ms_per_step": out,
            "learned_over_smagorinsky": round(out["learned"] / out["smagorinsky"], 3)}


def output_times(desc, tau, u_lid, reps=20):
    """Post-processing on the benchmark state after 100 steps: lbm_macroscopic (rho, u) and
    lbm_q_criterion, ms per call (HIP events on the launch stream) and algorithmic GB/s, where a
    node costs its populations + flag byte, plus its rho, u writes (macroscopic) or its own u
    read and Q write (Q-criterion; neighbour velocities are counted as cache hits)."""
    import torch

    from mojo_lbm_amd import _lib

    run = make_run(desc, u_lid)
    L, h, st = run["L"], run["h"], run["st"]
    _lib.check(L.lbm_set_option(h.ptr, _lib.OPT_STATIC_FLAGS, 1))  # the flags never change here
    time_run(run, tau, 100, 0)  # the state ends in run["f"]
    n, D, Q = desc.nx * desc.ny * desc.nz, desc.D, desc.Q
    rho = torch.empty(n, dtype=torch.float32, device="cuda")
    u = torch.empty(n * D, dtype=torch.float32, device="cuda")
    q = torch.empty(n, dtype=torch.float32, device="cuda")
    args = (run["f"].data_ptr(), run["bc"].data_ptr(), run["flags"].data_ptr())
    calls = {
        "macroscopic": (lambda: L.lbm_macroscopic(h.ptr, rho.data_ptr(), u.data_ptr(), *args, 1, _lib.U_XYZD, st),
                        n * (Q * 4 + 1 + 4 * (D + 1))),
        "q_criterion": (lambda: L.lbm_q_criterion(h.ptr, q.data_ptr(), u.data_ptr(), *args, tau, 1, _lib.U_XYZD, st),
                        n * (Q * 4 + 1 + 4 * D + 4)),
    }
    out = {}
    for name, (fn, nbytes) in calls.items():
        _lib.check(fn())
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        for _ in range(reps):
            fn()
        end.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(end) / reps
        out[name] = {"ms": round(ms, 5), "gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "roofline_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    del run
    torch.cuda.empty_cache()
    return {"workload": f"D3Q19 FP32 lid-driven cube {desc.nx}^3 after 100 steps", **out}


def load_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), if a profile summary
    for this kernel and workload is committed under profiles/."""
    p = ROOT / "profiles" / f"{name}.traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default=None,
                    help="run one variant: fp32 | f16c | f16c_ep | fp32_ep | fp32_periodic | f16c_periodic | cylinders | les "
                         "| outputs")
    ap.add_argument("--xcd-remap", type=int, default=None, help="override LBM_OPT_XCD_REMAP (0 or 1)")
    ap.add_argument("--staged-pull", type=int, default=None,
                    help="override LBM_OPT_STAGED_PULL (0 direct, 1 per tile, 2 per x-adjacent tile pair)")
    ap.add_argument("--step-lanes", type=int, default=None,
                    help="override LBM_OPT_STEP_LANES (lanes per step workgroup; 0 = one workgroup per tile)")
    ap.add_argument("--calibrate", action="store_true",
                    help="run only a plain 1 GiB -> 1 GiB device copy kernel (for scaling rocprofv3 byte counters)")
    ap.add_argument("--accuracy", default=None,
                    help="comma-separated checkpoints: run the Float16C / FP32 vs FP64 accuracy study instead")
    ap.add_argument("--cylinders", type=int, default=20000,
                    help="steps of the 64-case batched cylinder run reported under 'cylinder_batch' (0: skip)")
    ap.add_argument("--accuracy-steps", default="1000,5000,20000",
                    help="checkpoints of the Float16C / FP32 vs FP64 study folded into the default line ('' skips it)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip running every cylinder case alone after the batch")
    ap.add_argument("--cylinder-compare", action="store_true",
                    help="run only the batched cylinder cases, then every case alone, and compare")
    ap.add_argument("--cylinder-tile", default=None,
                    help="fixed tile shape for the cylinder batch, e.g. 32x8x1 (default: autotune)")
    args = ap.parse_args()
    if args.only == "les":
        print(json.dumps(les_step_times()), flush=True)
        return
    if args.only == "outputs":
        N, u_lid = args.n, 0.1
        from mojo_lbm_amd import _lib

        d = _lib.LbmDesc(D=3, Q=19, nx=N, ny=N, nz=N, tx=8, ty=8, tz=8, method=0, collision=0, storage=0,
                         ddf_shift=0, les=0, include_moving_boundary=1, included_bcs=3, smagorinsky_cs=0.1)
        print(json.dumps(output_times(d, 0.5 + 3.0 * u_lid * N / 100.0, u_lid)), flush=True)
        return
    if args.cylinder_compare or args.only == "cylinders":
        tile = tuple(int(v) for v in args.cylinder_tile.split("x")) if args.cylinder_tile else None
        print(json.dumps(cylinder_batch(args.cylinders or 20000, compare=args.cylinder_compare, tile=tile)),
              flush=True)
        return
    if args.accuracy:
        print(json.dumps(accuracy_study(args.n, [int(s) for s in args.accuracy.split(",")])), flush=True)
        return
    if args.calibrate:
        import torch

        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "torch.mul(x, 1.0, out=y) over 2^28 float32",
                          "read_bytes": 1 << 30, "write_bytes": 1 << 30}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (the sharded design is documented, not built). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    from mojo_lbm_amd import _lib
    from mojo_lbm_amd._lib import LbmDesc

    N = args.n
    u_lid, Re = 0.1, 100.0
    tau = 0.5 + 3.0 * u_lid * N / Re

    def desc_for(storage, method, ddf):
        return LbmDesc(D=3, Q=19, nx=N, ny=N, nz=N, tx=8, ty=8, tz=8, method=method, collision=0,
                       storage=storage, ddf_shift=ddf, les=0, include_moving_boundary=1, included_bcs=3,
                       smagorinsky_cs=0.1)

    tau_f16c = 0.5 + 3.0 * u_lid * N / 1000.0  # the Float16C case is quoted at Re = 1000
    variants = {
        "fp32": (0, 0, 0, 4, True, tau),      # storage, method, ddf, bytes per population, walls, tau
        "f16c": (1, 0, 1, 2, True, tau_f16c),
        "f16c_ep": (1, 1, 1, 2, True, tau_f16c),
        "fp32_ep": (0, 1, 0, 4, True, tau),
        # the reference's own benchmark box: no walls, fully periodic (kernels/benchmark/double_buffer.mojo:88-91)
        "fp32_periodic": (0, 0, 0, 4, False, tau),
        "f16c_periodic": (1, 0, 1, 2, False, tau_f16c),
    }
    names = [args.only] if args.only else list(variants)
    results = {}
    for name in names:
        storage, method, ddf, s, walls, tau_v = variants[name]
        d = desc_for(storage, method, ddf)
        run = make_run(d, u_lid, walls)
        # flags and bc are static in this benchmark: the flag summary is built once
        _lib.check(run["L"].lbm_set_option(run["h"].ptr, _lib.OPT_STATIC_FLAGS, 1))
        if args.xcd_remap is not None:
            _lib.check(run["L"].lbm_set_option(run["h"].ptr, _lib.OPT_XCD_REMAP, args.xcd_remap))
        if args.step_lanes is not None:
            _lib.check(run["L"].lbm_set_option(run["h"].ptr, _lib.OPT_STEP_LANES, args.step_lanes))
        if args.staged_pull is not None:
            _lib.check(run["L"].lbm_set_option(run["h"].ptr, _lib.OPT_STAGED_PULL, args.staged_pull))
        steps = args.steps if (name == "fp32" or args.only) else max(100, args.steps // 2)
        t = time_run(run, tau_v, steps, args.warmup)
        B = algorithmic_bytes(19, N, N, N, s, walls)
        ms = t / steps * 1e3
        results[name] = {
            "ms_per_step": ms,
            "tau": tau_v,
            "mlups": N ** 3 * steps / t / 1e6,
            "gbs": B / (t / steps) / 1e9,
            "bytes_per_step": B,
        }
        print(f"# {name}: {ms:.4f} ms/step, {results[name]['mlups']:.0f} MLUPs, "
              f"{results[name]['gbs']:.0f} GB/s algorithmic ({results[name]['gbs'] / HBM_PEAK_GBS:.1%} of 8 TB/s)",
              file=sys.stderr, flush=True)
        del run
        torch.cuda.empty_cache()

    head = names[0]
    r = results[head]
    traffic = load_traffic(f"step_d3q19_{head}_{N}")
    line = {
        "metric": "D3Q19 256^3 MLUPs & % HBM roofline (fp32/fp16c); learned-LES vs Smagorinsky",
        "value": round(r["mlups"], 1),
        "unit": "MLUPs",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if variants[head][0] == 0 else "f16c-storage/f32-math",
        "data": "synthetic (lid-driven cube initialised at equilibrium rho=1, u=0)",
        "config": {
            "workload": f"D3Q19 SRT lid-driven cube {N}^3, double-buffer pull, (x,y,z,q) 8x8x8 tiles, "
                        f"u_lid=0.1, Re=100, tau={tau:.3f}",
            "lattice": "D3Q19", "grid": [N, N, N], "tile": [8, 8, 8], "method": "double_buffer",
            "storage": head, "parallelism": "single GPU", "static_flags": True,
            "step_workgroups": ("FP32: one 512-lane workgroup per tile; Float16C: two 256-lane z slabs per tile "
                                "(the library default, LBM_OPT_STEP_LANES)") if args.step_lanes is None
                               else f"LBM_OPT_STEP_LANES = {args.step_lanes} for every variant",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(r["gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(r["gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "variants": {k: {"mlups": round(v["mlups"], 1), "ms_per_step": round(v["ms_per_step"], 5),
                         "roofline_frac": round(v["gbs"] / HBM_PEAK_GBS, 4), "tau": round(v["tau"], 4)}
                     for k, v in results.items()},
    }
    if "fp32" in results and "f16c" in results:
        line["variants"]["f16c_over_fp32"] = round(results["f16c"]["mlups"] / results["fp32"]["mlups"], 3)
    if args.cylinders and not args.only:
        # BL-cfg4: the 64 cases batched, then every case alone (populations and Cd / Cl compared)
        cyl = cylinder_batch(args.cylinders, compare=not args.no_compare)
        if "compare" in cyl:
            cmp_ = cyl["compare"]
            cmp_.pop("per_case", None)
            print(f"# cylinder batch vs single: {cmp_['bitwise_equal_cases']}/64 bitwise equal, "
                  f"max |dCd| {cmp_['max_abs_dcd']:.3g}, max |dCl| {cmp_['max_abs_dcl']:.3g}", file=sys.stderr, flush=True)
        print(f"# cylinder batch: {cyl['mlups']:.0f} MLUPs over {cyl['cases']} cases, tile {cyl['tile']}, "
              f"grid {cyl['grid_blocks']}, {cyl['roofline_frac']:.1%} of 8 TB/s", file=sys.stderr, flush=True)
        line["cylinder_batch"] = cyl
    if not args.only:
        outs = output_times(desc_for(0, 0, 0), tau, u_lid)
        print(f"# outputs: {outs}", file=sys.stderr, flush=True)
        line["outputs"] = outs
        les_t = les_step_times()
        print(f"# les 128^3: {les_t['ms_per_step']}", file=sys.stderr, flush=True)
        line["les_closure"] = les_t
        if args.accuracy_steps:
            # BL-cfg3: Float16C + DDF (DB and esoteric) and FP32 vs FP64 runs of the Re = 1000 cube
            line["accuracy"] = accuracy_study(N, [int(v) for v in args.accuracy_steps.split(",")])
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(desc_for(0, 0, 0), tau)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
