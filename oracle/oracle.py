"""ctypes wrapper of the CPU oracle (oracle/lbm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product path never uses it.
Arrays are numpy; f holds float32, float64 or uint16 Float16C codes, bc the grid
float type (float32 for F32/F16C, float64 for F64).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from mojo_lbm_amd._lib import LbmDesc

LIB_PATH = Path(__file__).resolve().parent / "_build" / "liblbm_oracle.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            from oracle.build import build

            build(verbose=False)
        L = C.CDLL(str(LIB_PATH))
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        pd = C.POINTER(LbmDesc)
        L.orc_set_threads.argtypes = [i]
        L.orc_set_q_clamp.argtypes = [C.c_float]
        L.orc_max_threads.restype = i
        L.orc_f16c_encode.argtypes = [C.c_float]
        L.orc_f16c_encode.restype = C.c_uint16
        L.orc_f16c_decode.argtypes = [C.c_uint16]
        L.orc_f16c_decode.restype = C.c_float
        L.orc_f16c_encode_n.argtypes = [vp, vp, C.c_size_t]
        L.orc_f16c_decode_n.argtypes = [vp, vp, C.c_size_t]
        L.orc_lattice.argtypes = [i, i, vp, vp, vp]
        L.orc_lattice.restype = i
        L.orc_addr.argtypes = [pd, i, i, i, i, i]
        L.orc_addr.restype = C.c_size_t
        L.orc_step_db.argtypes = [pd, vp, vp, vp, vp, d]
        L.orc_step_ep.argtypes = [pd, vp, vp, vp, d, i]
        L.orc_set_mlp.argtypes = [vp]
        L.orc_mlp_closure.argtypes = [vp, vp, C.c_long, vp]
        L.orc_step_db_links.argtypes = [pd, vp, vp, vp, vp, d, vp, vp, C.c_uint32, vp, vp]
        L.orc_step_db_links.restype = i
        L.orc_macroscopic_links.argtypes = [pd, vp, vp, vp, vp, vp, i, vp, vp, C.c_uint32]
        L.orc_macroscopic_links.restype = i
        L.orc_reduce_moments_links.argtypes = [pd, vp, vp, vp, vp, vp, vp, C.c_uint32]
        L.orc_reduce_moments_links.restype = i
        L.orc_macroscopic.argtypes = [pd, vp, vp, vp, vp, vp, i, i]
        L.orc_reduce_moments.argtypes = [pd, vp, vp, vp, vp, i]
        L.orc_init_equilibrium.argtypes = [pd, vp, d, vp]
        L.orc_set_exterior_wall.argtypes = [pd, vp, vp, i, i, i, vp, d]
        for n in ("orc_step_db", "orc_step_ep", "orc_macroscopic", "orc_reduce_moments",
                  "orc_init_equilibrium", "orc_set_exterior_wall"):
            getattr(L, n).restype = i
        L.orc_q_criterion.argtypes = [pd, vp, vp, i, vp, vp, vp, d, i]
        L.orc_q_criterion.restype = i
        L.orc_q_criterion_links.argtypes = [pd, vp, vp, i, vp, vp, vp, d, vp, vp, C.c_uint32]
        L.orc_q_criterion_links.restype = i
        L.orc_velocity_gradient.argtypes = [pd, vp, i, vp, i, i, i, i, i]
        L.orc_velocity_gradient.restype = d
        L.orc_init_f_from_fields.argtypes = [pd, vp, d, vp, vp, d]
        L.orc_init_f_from_fields.restype = i
        L.orc_set_threads(L.orc_max_threads())
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


def _ok(rc: int, what: str):
    if rc != 0:
        raise ValueError(f"oracle {what} rejected the description")


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def set_q_clamp(q: float) -> None:
    """Lower bound applied to every link fraction of the interpolated bounce-back (default 0)."""
    lib().orc_set_q_clamp(float(q))


def max_threads() -> int:
    return lib().orc_max_threads()


def f_dtype(desc: LbmDesc):
    return {0: np.float32, 1: np.uint16, 2: np.float64}[desc.storage]


def r_dtype(desc: LbmDesc):
    return np.float64 if desc.storage == 2 else np.float32


def n_cells(desc: LbmDesc) -> int:
    return desc.nx * desc.ny * desc.nz


def f16c_encode(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, np.uint16)
    lib().orc_f16c_encode_n(_p(x), _p(out), x.size)
    return out


def f16c_decode(h: np.ndarray) -> np.ndarray:
    h = np.ascontiguousarray(h, dtype=np.uint16)
    out = np.empty(h.shape, np.float32)
    lib().orc_f16c_decode_n(_p(h), _p(out), h.size)
    return out


def lattice(D: int, Q: int):
    c = np.zeros((Q, 3), np.int32)
    w = np.zeros(Q, np.float64)
    opp = np.zeros(Q, np.int32)
    _ok(lib().orc_lattice(D, Q, _p(c), _p(w), _p(opp)), "lattice")
    return c, w, opp


def addr(desc: LbmDesc, x: int, y: int, z: int, k: int, K: int) -> int:
    return lib().orc_addr(C.byref(desc), x, y, z, k, K)


_mlp_keep = None


def set_mlp(params) -> None:
    """Parameter block (1446 floats) of the learned closure used when desc.les == 2."""
    global _mlp_keep
    _mlp_keep = np.ascontiguousarray(params, np.float32)
    assert _mlp_keep.size == 1446
    lib().orc_set_mlp(_p(_mlp_keep))


def mlp_closure(params, X):
    """nu_t of the closure for raw features X [n, 9] (-1 outside the training range)."""
    P = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    out = np.zeros(len(X), np.float64)
    lib().orc_mlp_closure(_p(P), _p(X), len(X), _p(out))
    return out


def step_db(desc, f_out, f_in, bc, flags, tau):
    """One double-buffer step; with capture turned on in desc.included_bcs it writes bc in place."""
    _ok(lib().orc_step_db(C.byref(desc), _p(f_out), _p(f_in), _p(bc), _p(flags), float(tau)), "step_db")


def step_db_links(desc, f_out, f_in, bc, flags, tau, nodes, qdist, node_force=None):
    """Double-buffer step with interpolated bounce-back links; returns the momentum-exchange force (D,).
    node_force (float64 array of len(nodes) * D, optional) receives the force of each record's node."""
    nodes = np.ascontiguousarray(nodes, np.uint32)
    qdist = np.ascontiguousarray(qdist, np.float32)
    force = np.zeros(3, np.float64)
    if node_force is not None:
        assert node_force.dtype == np.float64 and node_force.size == len(nodes) * desc.D
    _ok(lib().orc_step_db_links(C.byref(desc), _p(f_out), _p(f_in), _p(bc), _p(flags), float(tau), _p(nodes),
                                _p(qdist), len(nodes), _p(force), None if node_force is None else _p(node_force)),
        "step_db_links")
    return force[: desc.D]


def step_ep(desc, f, bc, flags, tau, is_even):
    _ok(lib().orc_step_ep(C.byref(desc), _p(f), _p(bc), _p(flags), float(tau), int(bool(is_even))), "step_ep")


def _links(links):
    nodes = np.ascontiguousarray(links[0], np.uint32)
    qd = np.ascontiguousarray(links[1], np.float32)
    return nodes, qd


def macroscopic(desc, f, bc, flags, after_odd=False, u_layout=0, links=None):
    """rho, u after re-streaming; links = (nodes, qdist) applies interpolated bounce-back (double buffer)."""
    n = n_cells(desc)
    rho = np.zeros(n, r_dtype(desc))
    u = np.zeros(n * desc.D, r_dtype(desc))
    if links is not None:
        nodes, qd = _links(links)
        _ok(lib().orc_macroscopic_links(C.byref(desc), _p(rho), _p(u), _p(f), _p(bc), _p(flags), int(u_layout),
                                        _p(nodes), _p(qd), len(nodes)), "macroscopic_links")
        return rho, u
    _ok(lib().orc_macroscopic(C.byref(desc), _p(rho), _p(u), _p(f), _p(bc), _p(flags), int(after_odd),
                              int(u_layout)), "macroscopic")
    return rho, u


def q_criterion(desc, u, f, bc, flags, tau, after_odd=False, u_layout=0, links=None):
    """Q-criterion (row-major x, y, z) of the state; u is the velocity field in u_layout."""
    q = np.zeros(n_cells(desc), r_dtype(desc))
    u = np.ascontiguousarray(u, r_dtype(desc))
    if links is not None:
        nodes, qd = _links(links)
        _ok(lib().orc_q_criterion_links(C.byref(desc), _p(q), _p(u), int(u_layout), _p(f), _p(bc), _p(flags),
                                        float(tau), _p(nodes), _p(qd), len(nodes)), "q_criterion_links")
        return q
    _ok(lib().orc_q_criterion(C.byref(desc), _p(q), _p(u), int(u_layout), _p(f), _p(bc), _p(flags), float(tau),
                              int(after_odd)), "q_criterion")
    return q


def velocity_gradient(desc, u, flags, x, y, z, comp, axis, u_layout=0) -> float:
    """One finite-difference gradient d u_comp / d x_axis (FP64 velocity field)."""
    u = np.ascontiguousarray(u, np.float64)
    return lib().orc_velocity_gradient(C.byref(desc), _p(u), int(u_layout), _p(flags), x, y, z, comp, axis)


def reduce_moments(desc, f, bc, flags, after_odd=False, links=None):
    out = np.zeros(desc.D + 1, np.float64)
    if links is not None:
        nodes, qd = _links(links)
        _ok(lib().orc_reduce_moments_links(C.byref(desc), _p(out), _p(f), _p(bc), _p(flags), _p(nodes), _p(qd),
                                           len(nodes)), "reduce_links")
        return out
    _ok(lib().orc_reduce_moments(C.byref(desc), _p(out), _p(f), _p(bc), _p(flags), int(after_odd)), "reduce")
    return out


def init_equilibrium(desc, f, rho=1.0, u=None):
    ua = None if u is None else np.ascontiguousarray(u, np.float64)
    _ok(lib().orc_init_equilibrium(C.byref(desc), _p(f), float(rho), None if ua is None else _p(ua)), "init")


def init_f_from_fields(desc, f, rho, u, grad=None, tau=1.0):
    """initialize_f_from_func over a row-major velocity field (and gradient field) in the grid float type."""
    u = np.ascontiguousarray(u, r_dtype(desc))
    g = None if grad is None else np.ascontiguousarray(grad, r_dtype(desc))
    _ok(lib().orc_init_f_from_fields(C.byref(desc), _p(f), float(rho), _p(u), None if g is None else _p(g),
                                     float(tau)), "init_f_from_fields")


def set_exterior_wall(desc, flags, bc, side: str, flag: int, u=None, rho=float("nan")):
    axis = "XYZ".index(side[1])
    ua = None if u is None else np.ascontiguousarray(u, np.float64)
    _ok(lib().orc_set_exterior_wall(C.byref(desc), _p(flags), _p(bc), axis, int(side[0] == "+"), int(flag),
                                    None if ua is None else _p(ua), float(rho)), "set_exterior_wall")


def alloc(desc):
    """Zeroed f, bc (NaN-free zeros) and flags arrays for a description."""
    n = n_cells(desc)
    f = np.zeros(n * desc.Q, f_dtype(desc))
    bc = np.zeros(n * (desc.D + 1), r_dtype(desc))
    flags = np.zeros(n, np.uint8)
    return f, bc, flags
