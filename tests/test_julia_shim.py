"""Static checks of the Julia ccall shim (atomicorbitalkernels.jl_amd/julia/AOKHip.jl) against
the C ABI it binds (include/aok.h): there is no Julia toolchain here to run it, so every
``ccall((:aok_*, libaok), R, (T1, ...), a1, ...)`` is parsed and checked for

* a symbol that include/aok.h declares (and libaok.so exports, tests/test_abi.py),
* a return type and an argument-type tuple that match the C prototype position by position
  (aok_ctx / plans / graphs and data pointers -> Ptr{Cvoid}, enums and int -> Cint,
  unsigned -> Cuint, int64_t -> Int64, const int64_t * -> Ptr{Int64}, double -> Cdouble,
  out-handles -> Ref{Ptr{Cvoid}}, const char * -> Cstring),
* as many call arguments as declared types.

A mismatch here is the error a Julia caller would otherwise meet only at run time (or, for a
width mismatch, as silently wrong arguments)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "atomicorbitalkernels.jl_amd", "julia", "AOKHip.jl")
HEADER = os.path.join(ROOT, "include", "aok.h")

C_TO_JULIA = {
    "aok_ctx": "Ptr{Cvoid}", "aok_ovl2c_plan": "Ptr{Cvoid}", "aok_graph": "Ptr{Cvoid}",
    "aok_ctx*": "Ref{Ptr{Cvoid}}", "aok_ovl2c_plan*": "Ref{Ptr{Cvoid}}", "aok_graph*": "Ref{Ptr{Cvoid}}",
    "aok_dtype": "Cint", "aok_itype": "Cint", "aok_decay": "Cint", "aok_status": "Cint", "int": "Cint",
    "unsigned": "Cuint", "int64_t": "Int64", "double": "Cdouble",
    "void*": "Ptr{Cvoid}", "constvoid*": "Ptr{Cvoid}", "constint64_t*": "Ptr{Int64}",
    "constchar*": "Cstring", "int64_t*": "Ptr{Int64}", "double*": "Ptr{Cdouble}", "constchar**": "Ptr{Cstring}",
}


def _split_top(s):
    """Split on commas outside (), {} and []."""
    out, depth, cur = [], 0, []
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    if "".join(cur).strip():
        out.append("".join(cur).strip())
    return out


def _balanced(text, start):
    """Contents of the parenthesised group opening at text[start] == '('."""
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "(":
            depth += 1
        elif text[i] == ")":
            depth -= 1
            if depth == 0:
                return text[start + 1:i]
    raise ValueError("unbalanced parentheses")


def _c_prototypes():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    protos = {}
    for m in re.finditer(r"(aok_status|const char \*|double)\s*(aok_\w+)\s*\(([^;]*?)\)\s*;", src, flags=re.S):
        ret, name, params = m.group(1), m.group(2), m.group(3)
        types = []
        for p in _split_top(" ".join(params.split())):
            if p == "void":
                continue
            p = re.sub(r"\s*\b\w+$", "", p)  # drop the parameter name
            types.append(C_TO_JULIA.get(p.replace(" ", ""), f"<C {p}>"))
        protos[name] = (C_TO_JULIA[ret.replace(" ", "")], types)
    return protos


def _ccalls():
    text = open(SHIM).read()
    text = "\n".join(line.split("#", 1)[0] for line in text.splitlines())  # comments
    calls = []
    for m in re.finditer(r"ccall\(", text):
        body = _balanced(text, m.end() - 1)
        parts = _split_top(" ".join(body.split()))
        sym = re.match(r"\(:(\w+),\s*libaok\)", parts[0])
        assert sym, parts[0]
        ret = parts[1]
        assert parts[2].startswith("(") and parts[2].endswith(")"), parts[2]
        types = _split_top(parts[2][1:-1])
        calls.append((sym.group(1), ret, types, parts[3:], text[:m.start()].count("\n") + 1))
    return calls


def test_shim_has_ccalls():
    assert len(_ccalls()) >= 15


@pytest.mark.parametrize("call", _ccalls(), ids=lambda c: f"{c[0]}@{c[4]}")
def test_ccall_matches_header(call):
    name, ret, types, args, line = call
    protos = _c_prototypes()
    assert name in protos, f"AOKHip.jl:{line} binds {name}, which include/aok.h does not declare"
    cret, ctypes = protos[name]
    assert ret == cret, f"AOKHip.jl:{line} {name} returns {ret}, the header {cret}"
    assert types == ctypes, f"AOKHip.jl:{line} {name} argument types\n  shim   {types}\n  header {ctypes}"
    assert len(args) == len(types), f"AOKHip.jl:{line} {name}: {len(args)} arguments for {len(types)} types"


def test_every_shimmed_path_is_bound():
    """The entry points the shim exists for are all called from it."""
    bound = {c[0] for c in _ccalls()}
    for name in ("aok_create", "aok_ao_eval", "aok_ao_pullback", "aok_ovl2c", "aok_ovl2c_pullback",
                 "aok_ovl2c_pullback_ps", "aok_ovl2c_plan_create", "aok_ovl2c_plan_forward",
                 "aok_ovl2c_plan_pullback", "aok_ovl3c", "aok_ovl3c_pullback", "aok_ovl2c_pullback_x"):
        assert name in bound, name
