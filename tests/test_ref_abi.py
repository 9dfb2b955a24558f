"""The ctypes declarations of the reference-kernel harness (tests/ref_abi.py)
agree with its extern "C" prototypes (oracle/ref_harness.hip) parameter for
parameter.  Parses the source; needs neither the built harness nor a GPU."""
from __future__ import annotations

import ctypes as C
import re

import pytest
import ref_abi

C_TO_CTYPES = {
    "const char *": C.c_char_p,
    "int64_t": C.c_int64,
    "int32_t": C.c_int32,
    "int": C.c_int,
    "float": C.c_float,
    "double": C.c_double,
}


def prototypes():
    """name -> (return type, [parameter types]) of every function defined in
    the harness's extern "C" block."""
    text = open(ref_abi.SOURCE).read()
    text = re.sub(r"//[^\n]*", "", text)
    block = text[text.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"\n(const char \*|int|void)\s*([a-z_0-9]+)\s*\(([^)]*)\)\s*\{", block):
        ret, name, params = m.group(1), m.group(2), m.group(3)
        types = []
        for p in [p.strip() for p in params.split(",") if p.strip() and p.strip() != "void"]:
            p = re.sub(r"\s+", " ", p)
            ty = re.sub(r"\s*\b[a-z_0-9]+$", "", p).strip()   # drop the parameter name
            types.append(ty.replace(" *", "*").replace("*", " *").strip())
        out[name] = (ret, types)
    return out


def ctype_of(c_type):
    if c_type in C_TO_CTYPES:
        return C_TO_CTYPES[c_type]
    if c_type.endswith("*"):
        return C.c_void_p
    raise AssertionError(f"unmapped C type {c_type!r}")


def test_every_harness_function_is_declared():
    assert set(prototypes()) == set(ref_abi.SIGNATURES)


@pytest.mark.parametrize("name", sorted(ref_abi.SIGNATURES))
def test_declaration_matches_prototype(name):
    ret, params = prototypes()[name]
    res, args = ref_abi.SIGNATURES[name]
    assert ctype_of(ret) is res or (ret == "int" and res is C.c_int)
    assert len(args) == len(params), f"{name}: {len(args)} declared, {len(params)} in C"
    for k, (c_type, decl) in enumerate(zip(params, args)):
        assert ctype_of(c_type) is decl, f"{name} parameter {k}: C {c_type!r}, declared {decl}"
