"""On-disk snapshots of a run: checkpoint / resume for long runs and a store for
training data (no reference counterpart; the reference writes NumPy arrays of
derived fields only, src/utils/contextTileTensor.mojo:296-349).

Layout of one snapshot file:
  bytes 0..4095   UTF-8 JSON header padded with spaces: format name and version, the
                  run description (every lbm_desc field), step count, the esoteric-pull
                  parity, user metadata, and for each section (f, bc, flags) its byte
                  offset, byte length and CRC-32.
  sections        the raw device buffers in their tiled (x, y, z, k) layout, each at a
                  4096-byte aligned offset: f (storage type), bc (grid float type), flags.

Buffers move through pinned chunks, so a 512^3 Float16C state (5.1 GB of f) is
written and read without a second full-size host copy.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import zlib

import numpy as np

from . import _lib

MAGIC = "mojo_lbm_amd.snapshot"
VERSION = 1
HEADER_BYTES = 4096
CHUNK = 64 << 20
_DESC_FIELDS = [name for name, _ in _lib.LbmDesc._fields_]


def _align(n: int) -> int:
    return (n + HEADER_BYTES - 1) // HEADER_BYTES * HEADER_BYTES


def _desc_dict(desc) -> dict:
    return {k: getattr(desc, k) for k in _DESC_FIELDS}


def desc_from_dict(d: dict) -> _lib.LbmDesc:
    return _lib.LbmDesc(**{k: d[k] for k in _DESC_FIELDS})


def _as_bytes_view(t):
    """A flat uint8 view of a device or host tensor / array."""
    import torch

    if isinstance(t, np.ndarray):
        return torch.from_numpy(np.ascontiguousarray(t).view(np.uint8).reshape(-1))
    return t.contiguous().view(-1).view(torch.uint8)


def save(path, desc, f, bc, flags, step: int = 0, is_even: bool | None = None, meta: dict | None = None) -> dict:
    """Writes a snapshot; f, bc, flags are device tensors (or host arrays) in the library's layout.
    Returns the header."""
    import torch

    sections = {"f": _as_bytes_view(f), "bc": _as_bytes_view(bc), "flags": _as_bytes_view(flags)}
    expect = {"f": _lib.lib().lbm_f_bytes(C.byref(desc)), "bc": _lib.lib().lbm_bc_bytes(C.byref(desc)),
              "flags": _lib.lib().lbm_flags_bytes(C.byref(desc))}
    for k, v in sections.items():
        if v.numel() != expect[k]:
            raise ValueError(f"{k} holds {v.numel()} bytes, the description needs {expect[k]}")
    header = {"format": MAGIC, "version": VERSION, "desc": _desc_dict(desc), "step": int(step),
              "is_even": None if is_even is None else bool(is_even), "meta": meta or {}, "sections": {}}
    off = HEADER_BYTES
    for k, v in sections.items():
        header["sections"][k] = {"offset": off, "bytes": int(v.numel()), "crc32": 0}
        off = _align(off + int(v.numel()))
    pinned = torch.empty(CHUNK, dtype=torch.uint8, pin_memory=torch.cuda.is_available())
    # written beside the target and renamed over it once complete, so a crash mid-save leaves any
    # earlier snapshot at `path` intact
    tmp = f"{os.fspath(path)}.tmp"
    with open(tmp, "wb") as fh:
        fh.write(b" " * HEADER_BYTES)
        for k, v in sections.items():
            sec = header["sections"][k]
            fh.seek(sec["offset"])
            crc = 0
            for lo in range(0, sec["bytes"], CHUNK):
                n = min(CHUNK, sec["bytes"] - lo)
                pinned[:n].copy_(v[lo:lo + n])
                buf = pinned[:n].numpy().tobytes()
                crc = zlib.crc32(buf, crc)
                fh.write(buf)
            sec["crc32"] = crc
        text = json.dumps(header).encode()
        if len(text) > HEADER_BYTES:
            raise ValueError("snapshot metadata too large for the header")
        fh.seek(0)
        fh.write(text + b" " * (HEADER_BYTES - len(text)))
        fh.flush()
        os.fsync(fh.fileno())
    os.replace(tmp, path)
    return header


def read_header(path) -> dict:
    with open(path, "rb") as fh:
        raw = fh.read(HEADER_BYTES)
    try:
        header = json.loads(raw.decode().strip())
    except (UnicodeDecodeError, json.JSONDecodeError) as exc:
        raise ValueError(f"{path} is not a snapshot") from exc
    if header.get("format") != MAGIC:
        raise ValueError(f"{path} is not a snapshot")
    if header.get("version") != VERSION:
        raise ValueError(f"snapshot version {header.get('version')} is not supported (expected {VERSION})")
    return header


def load(path, f=None, bc=None, flags=None, device="cuda", verify: bool = True):
    """Reads a snapshot into the given tensors (allocated when None) on `device`.
    Returns (desc, f, bc, flags, header). With verify set, every section's CRC-32 is checked in
    a first pass over the file, before anything is copied, so a corrupt snapshot raises and
    leaves the destination tensors untouched."""
    import torch

    header = read_header(path)
    desc = desc_from_dict(header["desc"])
    outs = {"f": f, "bc": bc, "flags": flags}
    size = os.path.getsize(path)
    for k in ("f", "bc", "flags"):
        sec = header["sections"][k]
        if sec["offset"] + sec["bytes"] > size:
            raise ValueError(f"snapshot truncated in section {k}")
        if outs[k] is not None and _as_bytes_view(outs[k]).numel() != sec["bytes"]:
            raise ValueError(f"{k} holds {_as_bytes_view(outs[k]).numel()} bytes, the snapshot has {sec['bytes']}")
    if verify:
        with open(path, "rb") as fh:
            for k in ("f", "bc", "flags"):
                sec = header["sections"][k]
                fh.seek(sec["offset"])
                crc = 0
                for lo in range(0, sec["bytes"], CHUNK):
                    crc = zlib.crc32(fh.read(min(CHUNK, sec["bytes"] - lo)), crc)
                if crc != sec["crc32"]:
                    raise ValueError(f"snapshot section {k} fails its CRC-32 check")
    pinned = torch.empty(CHUNK, dtype=torch.uint8, pin_memory=torch.cuda.is_available())
    with open(path, "rb") as fh:
        for k in ("f", "bc", "flags"):
            sec = header["sections"][k]
            dst = outs[k]
            if dst is None:
                dst = torch.empty(sec["bytes"], dtype=torch.uint8, device=device)
                outs[k] = dst
            view = _as_bytes_view(dst)
            fh.seek(sec["offset"])
            for lo in range(0, sec["bytes"], CHUNK):
                n = min(CHUNK, sec["bytes"] - lo)
                buf = fh.read(n)
                pinned[:n].copy_(torch.frombuffer(bytearray(buf), dtype=torch.uint8))
                view[lo:lo + n].copy_(pinned[:n])
    return desc, outs["f"], outs["bc"], outs["flags"], header


def save_assembly(path, solver, assembly, step: int, which: str = "f", meta: dict | None = None) -> dict:
    """Snapshot of an Assembly driven by `solver`. `which` names the buffer holding the
    current populations ("f" after an odd double-buffer step or for esoteric pull, "f2"
    after an even double-buffer step). The next step is even when `step` is even."""
    src = getattr(assembly, which)
    if src is None:
        raise ValueError(f"the assembly has no {which} buffer")
    return save(path, solver.desc, src.gpu(), assembly.bc.gpu(), assembly.flags.gpu(), step=step,
                is_even=(step % 2 == 0), meta=meta)


def load_assembly(path, solver, assembly) -> dict:
    """Restores an Assembly from a snapshot written for the same description. The
    populations land where the next step reads them: for double buffer in f when the
    header's step is even (even_step reads f) and in f2 when it is odd; for esoteric
    pull in f. Resume with even_step when the step is even. Returns the header."""
    header = read_header(path)
    ours = _desc_dict(solver.desc)
    if header["desc"] != ours:
        diff = sorted(k for k in ours if header["desc"].get(k) != ours[k])
        raise ValueError(f"snapshot description differs from the solver's in {diff}")
    dst = assembly.f if (assembly.f2 is None or header["step"] % 2 == 0) else assembly.f2
    load(path, dst.gpu(), assembly.bc.gpu(), assembly.flags.gpu())
    for t in (dst, assembly.bc, assembly.flags):
        t.mark_device_written()
    return header
