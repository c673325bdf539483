"""The Python mirror of the C ABI agrees with include/lbm_hip.h: every enum constant of the
header has the same value in mojo_lbm_amd._lib (or, for node flags, mojo_lbm_amd.constants.Flags),
and the ctypes lbm_desc has the header's fields, in order, with the header's types."""
import ctypes as C
import re
from pathlib import Path

from mojo_lbm_amd import _lib
from mojo_lbm_amd.constants import Flags

HEADER = Path(__file__).resolve().parents[1] / "include" / "lbm_hip.h"


def _strip_comments(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def _enum_constants():
    text = _strip_comments(HEADER.read_text())
    out = {}
    for body in re.findall(r"enum\s*\{(.*?)\}", text, flags=re.S):
        for name, value in re.findall(r"(LBM_\w+)\s*=\s*(-?(?:0x[0-9A-Fa-f]+|\d+))", body):
            out[name] = int(value, 0)
    return out


def test_every_header_constant_is_mirrored():
    consts = _enum_constants()
    assert len(consts) >= 30
    for name, value in consts.items():
        short = name[len("LBM_"):]
        if short.startswith("FLAG_"):
            assert getattr(Flags, short[len("FLAG_"):]) == value, name
        else:
            assert hasattr(_lib, short), f"{name} has no counterpart in mojo_lbm_amd._lib"
            assert getattr(_lib, short) == value, name


def test_desc_layout_matches_header():
    text = _strip_comments(HEADER.read_text())
    body = re.search(r"typedef struct lbm_desc\s*\{(.*?)\}\s*lbm_desc;", text, flags=re.S).group(1)
    ctype = {"int32_t": C.c_int32, "uint32_t": C.c_uint32, "double": C.c_double}
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        typ, names = decl.split(None, 1)
        fields += [(n.strip(), ctype[typ]) for n in names.split(",")]
    assert [(n, t) for n, t in _lib.LbmDesc._fields_] == fields
