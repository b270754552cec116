"""The two host layers against the C ABI they bind.  The Julia boundary
(julia/src/BoundaryIntegralEquationsAMD.jl; Julia is not installed here): every `ccall((:name, libbie[]), R, (T1, ..., Tn), ...)`
names a function that include/bie_amd.h declares and the shared library exports, with as
many argument types as the declaration has parameters, and a Cint result where the
declaration returns int.  The Python layer (bie_amd._lib.SIGNATURES): one ctypes signature
per declared function, with the declared number of arguments.  A renamed entry point or an
added parameter that either side missed fails here.
"""
import ctypes
import os
import re

from bie_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "boundaryintegralequations.jl_amd", "julia", "src",
                  "BoundaryIntegralEquationsAMD.jl")
HEADER = os.path.join(ROOT, "include", "bie_amd.h")


def _group(text, start, open_ch="(", close_ch=")"):
    """The text between text[start] == open_ch and its matching close_ch."""
    depth = 0
    for i in range(start, len(text)):
        if text[i] == open_ch:
            depth += 1
        elif text[i] == close_ch:
            depth -= 1
            if depth == 0:
                return text[start + 1:i]
    raise ValueError("unbalanced")


def _count_top_level(items):
    items = items.strip().rstrip(",").strip()   # Julia's one-element tuple (T,)
    if not items or items == "void":
        return 0
    depth, n = 0, 1
    for ch in items:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        elif ch == "," and depth == 0:
            n += 1
    return n


def _header_decls():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    decls = {}
    for m in re.finditer(r"(?m)^\s*([A-Za-z_][\w \*]*?)\b(bie_\w+)\s*\(", text):
        ret = m.group(1).strip()
        params = _group(text, m.end() - 1)
        decls[m.group(2)] = (ret, _count_top_level(params))
    return decls


def _julia_calls():
    text = open(JL).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(\w+),\s*libbie\[\]\),\s*(\w+),\s*", text):
        types = _group(text, m.end())
        calls.append((m.group(1), m.group(2), _count_top_level(types)))
    return calls


def test_julia_ccalls_match_the_header_and_the_library():
    decls = _header_decls()
    calls = _julia_calls()
    assert len(calls) >= 20
    lib = _lib.lib()
    for name, ret, nargs in calls:
        assert name in decls, f"{name} is not declared in include/bie_amd.h"
        hret, hargs = decls[name]
        assert nargs == hargs, f"{name}: Julia passes {nargs} types, the header has {hargs}"
        if hret == "int":
            assert ret == "Cint", f"{name} returns int"
        assert hasattr(lib, name) and isinstance(getattr(lib, name), ctypes._CFuncPtr)
    # the near field with Ngauss is what the Julia hook calls
    assert any(n == "bie_sparse_assemble_ex" for n, _, _ in calls)


def test_python_signatures_match_the_header():
    """The ctypes table of the Python host layer (bie_amd._lib.SIGNATURES) against the
    header: every entry is declared, with as many argument types as parameters, and every
    function the header declares has an entry (so Python can reach all of them)."""
    decls = _header_decls()
    for name, (restype, argtypes) in _lib.SIGNATURES.items():
        assert name in decls, f"{name} is not declared in include/bie_amd.h"
        assert len(argtypes) == decls[name][1], f"{name}: {len(argtypes)} ctypes arguments"
        if decls[name][0] == "int":
            assert restype is ctypes.c_int, f"{name} returns int"
    missing = sorted(set(decls) - set(_lib.SIGNATURES))
    assert not missing, f"declared but without a ctypes signature: {missing}"
