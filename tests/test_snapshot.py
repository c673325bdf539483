"""Snapshot files: round trip of the description and every buffer, CRC and version checks
(CPU), and checkpoint / resume reproducing an uninterrupted run bit for bit (GPU)."""
import json

import numpy as np
import pytest

from mojo_lbm_amd import snapshot
from tests import _util as U


def _state(desc, seed=0):
    f, bc, flags = U.perturbed_state(desc, seed=seed)
    return f, bc, flags


def test_round_trip_on_the_host(tmp_path):
    desc = U.make_desc("D3Q19", (16, 16, 16), storage=1, ddf=1)
    f, bc, flags = _state(desc)
    p = tmp_path / "s.lbm"
    hdr = snapshot.save(p, desc, f, bc, flags, step=1234, is_even=False, meta={"note": "x"})
    f2, bc2, fl2 = np.zeros_like(f), np.zeros_like(bc), np.zeros_like(flags)
    d2, _, _, _, h2 = snapshot.load(p, f2, bc2, fl2, device="cpu")
    assert np.array_equal(f, f2) and np.array_equal(bc, bc2) and np.array_equal(flags, fl2)
    assert h2["step"] == 1234 and h2["is_even"] is False and h2["meta"] == {"note": "x"}
    assert all(getattr(d2, k) == getattr(desc, k) for k in snapshot._DESC_FIELDS)
    assert hdr["sections"]["f"]["offset"] % 4096 == 0 and hdr["sections"]["flags"]["offset"] % 4096 == 0


def test_save_replaces_the_file_only_when_complete(tmp_path, monkeypatch):
    """A save that fails part way leaves the earlier snapshot at the path readable."""
    desc = U.make_desc("D2Q9", (32, 32))
    f, bc, flags = _state(desc, 2)
    p = tmp_path / "s.lbm"
    snapshot.save(p, desc, f, bc, flags, step=10)
    real = snapshot.zlib.crc32
    calls = {"n": 0}

    def failing_crc(buf, crc=0):
        calls["n"] += 1
        if calls["n"] == 2:
            raise OSError("disk full")
        return real(buf, crc)

    monkeypatch.setattr(snapshot.zlib, "crc32", failing_crc)
    with pytest.raises(OSError):
        snapshot.save(p, desc, f + 1, bc, flags, step=20)
    monkeypatch.setattr(snapshot.zlib, "crc32", real)
    f2, bc2, fl2 = np.zeros_like(f), np.zeros_like(bc), np.zeros_like(flags)
    _, _, _, _, h = snapshot.load(p, f2, bc2, fl2, device="cpu")
    assert h["step"] == 10 and np.array_equal(f2, f)


def test_corruption_and_version_are_detected(tmp_path):
    desc = U.make_desc("D2Q9", (32, 32))
    f, bc, flags = _state(desc, 1)
    p = tmp_path / "s.lbm"
    hdr = snapshot.save(p, desc, f, bc, flags)
    raw = bytearray(p.read_bytes())
    raw[hdr["sections"]["bc"]["offset"] + 5] ^= 0x10
    p.write_bytes(bytes(raw))
    dst = (np.full_like(f, 7), np.full_like(bc, 7), np.full_like(flags, 7))
    with pytest.raises(ValueError, match="CRC"):
        snapshot.load(p, *dst, device="cpu")
    # the corrupt section is the second one: nothing, not even the intact first section, was copied
    assert np.all(dst[0] == 7) and np.all(dst[1] == 7) and np.all(dst[2] == 7)
    h = json.loads(raw[:4096].decode())
    h["version"] = 99
    text = json.dumps(h).encode()
    raw[:4096] = text + b" " * (4096 - len(text))
    p.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="version"):
        snapshot.read_header(p)
    with pytest.raises(ValueError, match="needs"):
        snapshot.save(tmp_path / "t.lbm", desc, f[:-1], bc, flags)


@pytest.mark.gpu
@pytest.mark.parametrize("method", [0, 1])
def test_checkpoint_resume_is_exact(gpu, tmp_path, method):
    import torch

    desc = U.make_desc("D3Q19", (32, 32, 32), method=method, storage=1, ddf=1)
    f, bc, flags = _state(desc, 2)
    run = U.DeviceRun(desc, f, bc, flags)
    step = (lambda k: run.step_db(0.6)) if method == 0 else (lambda k: run.step_ep(0.6, k % 2 == 0))
    for k in range(7):
        step(k)
    p = tmp_path / "ck.lbm"
    snapshot.save(p, desc, run.f, run.bc, run.flags, step=7, is_even=(7 % 2 == 0))
    for k in range(7, 12):
        step(k)
    ref = run.f_host()
    d2, f2, bc2, fl2, hdr = snapshot.load(p)
    run2 = U.DeviceRun(d2, f2.cpu().numpy().view(f.dtype), bc2.cpu().numpy().view(bc.dtype), fl2.cpu().numpy())
    step2 = (lambda k: run2.step_db(0.6)) if method == 0 else (lambda k: run2.step_ep(0.6, k % 2 == 0))
    for k in range(hdr["step"], 12):
        step2(k)
    assert np.array_equal(run2.f_host(), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("method_name", ["DOUBLE_BUFFER", "ESOTERIC_PULL"])
def test_assembly_checkpoint_resume(gpu, tmp_path, method_name):
    """A D2Q9 lid-driven cavity through the host API: 9 steps, save, 5 more steps; a
    fresh Assembly restored from the file and stepped 5 times ends in the same state."""
    import mojo_lbm_amd as lbm

    method = getattr(lbm, method_name)

    def setup():
        grid = lbm.LBM_Grid(lbm.get_D2Q9(), 64, 64, 1, 16)
        config = lbm.LBM_Config(method, use_float16c=True, DDF_shift=True)
        solver = lbm.Solver(grid, config)
        asm = lbm.Assembly(grid, config)
        lbm.initialize_fluid_at_rest(grid, config, asm.f, solver)
        if asm.f2 is not None:
            lbm.initialize_fluid_at_rest(grid, config, asm.f2, solver)
        lbm.set_exterior_walls(grid, config, asm.flags, asm.bc, "+Y", lbm.Flags.SOLID, [0.1, 0.0], 1.0)
        for side in ("-Y", "+X", "-X"):
            lbm.set_exterior_walls(grid, config, asm.flags, asm.bc, side, lbm.Flags.SOLID, [0.0, 0.0], 1.0)
        return solver, asm

    def step(solver, asm, k):
        (solver.even_step if k % 2 == 0 else solver.odd_step)(asm, 0.56)

    def current(asm, k):  # buffer holding the populations after k steps
        return asm.f2 if (asm.f2 is not None and k % 2 == 1) else asm.f

    solver, asm = setup()
    for k in range(9):
        step(solver, asm, k)
    p = tmp_path / "cavity.lbm"
    snapshot.save_assembly(p, solver, asm, 9, which="f2" if method_name == "DOUBLE_BUFFER" else "f")
    for k in range(9, 14):
        step(solver, asm, k)
    ref = current(asm, 14).cpu().copy()

    solver2, asm2 = setup()
    hdr = snapshot.load_assembly(p, solver2, asm2)
    assert hdr["step"] == 9
    for k in range(9, 14):
        step(solver2, asm2, k)
    assert np.array_equal(current(asm2, 14).cpu(), ref)


@pytest.mark.gpu
def test_check_health_flags_divergence(gpu):
    """check_health: finite totals pass, a mass drift beyond the bound or a NaN population raises."""
    import torch

    import mojo_lbm_amd as lbm

    grid = lbm.LBM_Grid(lbm.get_D2Q9(), 32, 32, 1, 16)
    config = lbm.LBM_Config(lbm.DOUBLE_BUFFER)
    solver = lbm.Solver(grid, config)
    asm = lbm.Assembly(grid, config)
    lbm.initialize_fluid_at_rest(grid, config, asm.f, solver)
    lbm.initialize_fluid_at_rest(grid, config, asm.f2, solver)
    lbm.set_exterior_walls(grid, config, asm.flags, asm.bc, "+Y", lbm.Flags.SOLID, [0.1, 0.0], 1.0)
    for k in range(4):
        (solver.even_step if k % 2 == 0 else solver.odd_step)(asm, 0.7)
    tot = lbm.check_health(solver, asm, True)
    lbm.check_health(solver, asm, True, mass0=float(tot[0]))
    with pytest.raises(FloatingPointError, match="drifted"):
        lbm.check_health(solver, asm, True, mass0=float(tot[0]) * 1.01)
    f = asm.f.gpu().view(torch.float32)
    f[123] = float("nan")
    asm.f.mark_device_written()
    with pytest.raises(FloatingPointError, match="diverged"):
        lbm.check_health(solver, asm, True)
