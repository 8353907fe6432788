"""The ctypes binding's argument and return types against the C prototypes.

test_abi.py checks that every declared symbol is exported and bound by name; this file
parses each prototype of include/fft_cffi.h and include/ft_tools_cffi.h and checks that
the binding (fftfree_amd/_lib.py) passes every argument as the C type the library reads:
an int bound as size_t, a float bound as int or a missing argument would otherwise shift
or truncate arguments silently under the SysV calling convention (the failure the
reference's own Python cdefs have, fft_to_png.py:43-93 vs fft_cffi.hpp:99-150). CPU only.
"""
import ctypes
import re
from pathlib import Path

import pytest

from fftfree_amd import _lib

INCLUDE = Path(__file__).resolve().parents[1] / "include"
HEADERS = [INCLUDE / "fft_cffi.h", INCLUDE / "ft_tools_cffi.h"]

# C scalar spellings -> the ctypes type that must bind them
SCALARS = {
    "int": ctypes.c_int,
    "unsigned int": ctypes.c_uint,
    "unsigned": ctypes.c_uint,
    "size_t": ctypes.c_size_t,
    "float": ctypes.c_float,
    "long long": ctypes.c_longlong,
}
POINTER_TYPEDEFS = {"fft_restore_observer_t"}


def prototypes():
    """name -> (return C type, [parameter C types]) for every function declared."""
    out = {}
    for header in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        text = re.sub(r"typedef[^;]*;", "", text)  # typedefs, including the ragged request struct
        for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b((?:fft|phase|ft)_\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
            ret, name, params = m.group(1), m.group(2), m.group(3)
            ret = " ".join(ret.split())
            plist = []
            for p in [p.strip() for p in params.split(",")]:
                if p in ("", "void"):
                    continue
                p = " ".join(p.replace("*", " * ").split())
                # drop the parameter name (last identifier, when the declaration has one)
                toks = p.split(" ")
                if len(toks) > 1 and toks[-1] not in ("*",) and re.fullmatch(r"\w+", toks[-1]) and \
                        " ".join(toks[:-1]).replace("const ", "").strip() not in ("",):
                    p = " ".join(toks[:-1])
                plist.append(p)
            out[name] = (ret, plist)
    return out


def expected(ctype: str):
    """('ptr', None) for pointers, ('scalar', ctypes type) for scalars, ('void', None)."""
    t = ctype.replace("const ", "").replace("struct ", "").strip()
    if t == "void":
        return ("void", None)
    if "*" in t or t in POINTER_TYPEDEFS:
        return ("ptr", t)
    if t not in SCALARS:
        raise AssertionError(f"unexpected C type {ctype!r}")
    return ("scalar", SCALARS[t])


def bound_kind(ct):
    if ct is None:
        return ("void", None)
    if ct in (ctypes.c_void_p, ctypes.c_char_p) or (isinstance(ct, type) and issubclass(ct, ctypes._Pointer)):
        return ("ptr", ct)
    return ("scalar", ct)


def pointer_ok(c_spelling: str, ct) -> bool:
    """A pointer argument may be bound as void* or as a pointer to the same element."""
    if ct is ctypes.c_void_p:
        return True
    base = c_spelling.replace("*", "").strip()
    if ct is ctypes.c_char_p:
        return base == "char"
    elem = ct._type_
    if base == "float":
        return elem is ctypes.c_float
    if base == "int":
        return elem is ctypes.c_int
    if base == "fft_ragged_request":
        return elem is _lib.RaggedRequest
    return False


PROTOS = prototypes()


def test_every_bound_function_has_a_parsed_prototype():
    assert set(PROTOS) == set(_lib.SIGNATURES), set(PROTOS) ^ set(_lib.SIGNATURES)


@pytest.mark.parametrize("name", sorted(_lib.SIGNATURES))
def test_binding_types_match_the_prototype(name):
    ret_c, params_c = PROTOS[name]
    res, args = _lib.SIGNATURES[name]
    assert len(args) == len(params_c), f"{name}: {len(args)} bound arguments, prototype has {len(params_c)}"
    kind, want = expected(ret_c)
    got = bound_kind(res)
    assert kind == got[0], f"{name}: return {ret_c} bound as {res}"
    if kind == "scalar":
        assert res is want, f"{name}: return {ret_c} bound as {res}"
    for i, (pc, ct) in enumerate(zip(params_c, args)):
        kind, want = expected(pc)
        got = bound_kind(ct)
        assert kind == got[0], f"{name} argument {i}: {pc} bound as {ct}"
        if kind == "scalar":
            assert ct is want, f"{name} argument {i}: {pc} bound as {ct}"
        else:
            assert pointer_ok(want, ct), f"{name} argument {i}: {pc} bound as {ct}"


def test_checker_catches_a_slipped_argument():
    """Planted mistakes in a copy of the table must fail: an int bound as size_t, a
    float bound as int, and a dropped argument."""
    ret_c, params_c = PROTOS["fft_init_full_v2"]
    res, args = _lib.SIGNATURES["fft_init_full_v2"]

    def mismatches(bound):
        bad = 0
        if len(bound) != len(params_c):
            return 1
        for pc, ct in zip(params_c, bound):
            kind, want = expected(pc)
            if bound_kind(ct)[0] != kind or (kind == "scalar" and ct is not want):
                bad += 1
        return bad

    assert mismatches(args) == 0
    slipped = list(args)
    slipped[9] = ctypes.c_size_t  # `int window` as size_t
    assert mismatches(slipped) == 1
    slipped = list(args)
    slipped[23] = ctypes.c_int  # `float analysis_param1` as int
    assert mismatches(slipped) == 1
    assert mismatches(args[:-1]) == 1
