"""The C-ABI library: it loads, exports every function include/lbm_hip.h declares,
and rejects invalid run descriptions the way the reference's comptime asserts and
host checks do. No compute call is made here (no GPU in this suite)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from mojo_lbm_amd import _lib
from tests import _util as U

HEADER = Path(__file__).resolve().parents[1] / "include" / "lbm_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/lbm_hip.h but not exported"
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_desc_struct_layout_matches_header():
    # 14 int32 + uint32 + (4 pad) + double
    assert C.sizeof(_lib.LbmDesc) == 72
    assert _lib.LbmDesc.smagorinsky_cs.offset == 64


def test_version_and_device_count():
    assert _lib.lib().lbm_version() >= 100
    assert _lib.device_count() >= 0


def _sizes(desc):
    L = _lib.lib()
    return L.lbm_f_bytes(C.byref(desc)), L.lbm_bc_bytes(C.byref(desc)), L.lbm_flags_bytes(C.byref(desc))


def test_buffer_sizes():
    d = U.make_desc("D3Q19", (32, 16, 8), storage=0)
    assert _sizes(d) == (32 * 16 * 8 * 19 * 4, 32 * 16 * 8 * 4 * 4, 32 * 16 * 8)
    d = U.make_desc("D3Q19", (32, 16, 8), storage=1)
    assert _sizes(d)[0] == 32 * 16 * 8 * 19 * 2
    d = U.make_desc("D2Q9", (64, 32, 1), storage=2)
    assert _sizes(d) == (64 * 32 * 9 * 8, 64 * 32 * 3 * 8, 64 * 32)


def _create_rc(desc):
    h = C.c_void_p()
    rc = _lib.lib().lbm_create(C.byref(desc), C.byref(h))
    if rc == 0:
        _lib.lib().lbm_destroy(h)
    return rc, _lib.lib().lbm_last_error().decode()


@pytest.mark.parametrize(
    "mod,msg",
    [
        (dict(Q=15), "lattice"),
        (dict(nx=20), "divisible"),  # grid.mojo:419-440
        (dict(tx=6), "powers of two"),
        (dict(tx=16, nx=32), "workgroup"),  # 16x8x8 nodes per tile > 512
        (dict(method=3), "method"),
        (dict(storage=7), "storage"),
        (dict(les=5), "LES"),
        (dict(smagorinsky_cs=-1.0), "smagorinsky"),
    ],
)
def test_create_rejects_invalid_descriptions(mod, msg):
    d = U.make_desc_like(U.make_desc("D3Q19", (32, 16, 16)), **mod)
    rc, err = _create_rc(d)
    assert rc == _lib.ERR_INVALID
    assert msg.lower() in err.lower()


def test_collision_operators_validate():
    # every operator of the reference's Collisions set passes validation (the CPU box then reports
    # no device); an unknown id is invalid
    for coll in (_lib.COLL_SRT, _lib.COLL_TRT, _lib.COLL_KBC, _lib.COLL_RLBM):
        rc, err = _create_rc(U.make_desc("D2Q9", (32, 32), collision=coll))
        assert rc in (_lib.OK, _lib.ERR_NO_DEVICE), err
    rc, err = _create_rc(U.make_desc("D2Q9", (32, 32), collision=7))
    assert rc == _lib.ERR_INVALID


def test_learned_closure_needs_3d():
    rc, err = _create_rc(U.make_desc("D2Q9", (32, 32), les=2))
    assert rc == _lib.ERR_UNSUPPORTED, err


def test_create_without_gpu_reports_no_device():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    rc, err = _create_rc(U.make_desc("D3Q19", (16, 16, 16)))
    assert rc == _lib.ERR_NO_DEVICE


def test_calls_with_null_handle_fail_cleanly():
    L = _lib.lib()
    assert L.lbm_step_db(None, None, None, None, None, 1.0, None) == _lib.ERR_INVALID
    assert L.lbm_step_ep(None, None, None, None, 1.0, 1, None) == _lib.ERR_INVALID
    assert L.lbm_destroy(None) == 0
