"""Write abi.jl (the Julia ccall layer of libexactclosure) and the ccall table
of INTEGRATION.md from include/exactclosure.h and the ctypes declarations of
_abi.py.

One Julia function per C entry point, named like the C function and taking
the C parameter names; the argument types are the ctypes types of _abi.py
mapped to Julia's (so the Python and Julia bindings cannot disagree).

    python exactclosure.jl_amd/julia/gen_abi.py          # rewrite both
    python exactclosure.jl_amd/julia/gen_abi.py --check  # exit 1 if stale
"""
from __future__ import annotations

import ctypes
import importlib.util
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
HEADER = os.path.join(ROOT, "include", "exactclosure.h")
ABI_JL = os.path.join(HERE, "abi.jl")
INTEGRATION = os.path.join(ROOT, "INTEGRATION.md")
TABLE_BEGIN = "<!-- ccall table: written by exactclosure.jl_amd/julia/gen_abi.py -->"
TABLE_END = "<!-- end of ccall table -->"

# Julia reserves these words; a C parameter of that name gets a trailing underscore.
JULIA_WORDS = {"in", "end", "function", "type", "begin", "do", "let", "local", "global", "const", "module",
               "struct", "for", "while", "if", "else", "try", "catch", "return", "break", "continue", "quote",
               "macro", "import", "using", "export", "isa", "where", "true", "false", "mutable", "abstract"}


def load_abi():
    spec = importlib.util.spec_from_file_location("_ec_abi", os.path.join(PKG, "_abi.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def header_params(path=HEADER):
    """{name: [(parameter name, C declaration)]} of every function declared in
    the header."""
    txt = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    out = {}
    for m in re.finditer(r"(?:^|\n)\s*(?:int|const char \*|void \*)\s*(ec_\w+)\s*\(([^;]*?)\)\s*;", txt):
        name, args = m.group(1), " ".join(m.group(2).split())
        names = []
        if args not in ("", "void"):
            for a in args.split(","):
                a = re.sub(r"\[\w*\]", "", a.strip())
                pname = re.findall(r"(\w+)\s*$", a)[0]
                names.append((pname + "_" if pname in JULIA_WORDS else pname, a))
        out[name] = names
    return out


def julia_type(t, abi):
    """The Julia ccall type of a ctypes type of _abi.py."""
    # (ctypes.c_size_t is ctypes.c_uint64 here: size_t parameters are told
    # apart by their C declaration in entries())
    simple = {ctypes.c_int: "Cint", ctypes.c_double: "Cdouble", ctypes.c_uint64: "UInt64",
              ctypes.c_float: "Cfloat", ctypes.c_void_p: "Ptr{Cvoid}", ctypes.c_char_p: "Cstring"}
    if t in simple:
        return simple[t]
    structs = {abi.FilterQuantity: "FilterQuantity", abi.VelSet: "VelSet", abi.ClosureSet: "ClosureSet"}
    if t in structs:
        return structs[t]
    if isinstance(t, type) and issubclass(t, ctypes.Array):
        # void *const x[k] on the host: a pointer to k device pointers
        return f"Ref{{NTuple{{{t._length_},{julia_type(t._type_, abi)}}}}}"
    if isinstance(t, type) and issubclass(t, ctypes._Pointer):
        inner = t._type_
        if inner is ctypes.c_char_p:
            return "Ptr{Cstring}"
        return f"Ptr{{{julia_type(inner, abi)}}}"
    raise TypeError(f"no Julia type for {t!r}")


def julia_restype(t, abi):
    return "Cvoid" if t is None else julia_type(t, abi)


def struct_field_type(t, abi):
    if isinstance(t, type) and issubclass(t, ctypes.Array):
        return f"NTuple{{{t._length_},{julia_type(t._type_, abi)}}}"
    return julia_type(t, abi)


def entries(abi=None):
    """[(name, restype, [(param, type)])] in header order."""
    abi = abi or load_abi()
    params = header_params()
    out = []
    for name, pnames in params.items():
        if name not in abi._SIGS:
            raise KeyError(f"{name} is declared in the header but not in _abi.py")
        res, args = abi._SIGS[name]
        if len(args) != len(pnames):
            raise ValueError(f"{name}: header has {len(pnames)} parameters, _abi.py {len(args)}")
        jargs = []
        for (p, decl), a in zip(pnames, args):
            jt = julia_type(a, abi)
            if re.match(r"(const\s+)?size_t\b", decl):
                jt = "Csize_t"
            jargs.append((p, jt))
        out.append((name, julia_restype(res, abi), jargs))
    return out


def render_abi_jl(abi=None):
    abi = abi or load_abi()
    lines = [
        "# The ccall layer of libexactclosure: one Julia function per C entry point of",
        "# include/exactclosure.h, same name, C parameter names, argument types as in",
        "# the header.  Written by gen_abi.py from the header and _abi.py; edit those,",
        "# not this file.  Included by ExactClosureAMD.jl, which defines",
        "# `libexactclosure` (the path of the shared library).",
        "",
    ]
    for sname, cls in (("FilterQuantity", abi.FilterQuantity), ("VelSet", abi.VelSet),
                       ("ClosureSet", abi.ClosureSet)):
        cname = {"FilterQuantity": "ec_filter_quantity", "VelSet": "ec_vel_set", "ClosureSet": "ec_closure_set"}[sname]
        lines.append(f"# {cname}")
        lines.append(f"struct {sname}")
        for fname, ftype in cls._fields_:
            lines.append(f"    {fname}::{struct_field_type(ftype, abi)}")
        lines.append("end")
        lines.append("")
    for name, res, args in entries(abi):
        pn = ", ".join(p for p, _ in args)
        tys = ", ".join(t for _, t in args)
        tup = f"({tys},)" if len(args) == 1 else f"({tys})"
        call = f"ccall((:{name}, libexactclosure), {res}, {tup}" + (f", {pn})" if args else ")")
        lines.append(f"{name}({pn}) =")
        lines.append(f"    {call}")
    return "\n".join(lines) + "\n"


def render_table(abi=None):
    rows = [TABLE_BEGIN, "", "| symbol | returns | argument types |", "|---|---|---|"]
    for name, res, args in entries(abi):
        tys = ", ".join(t for _, t in args)
        tup = f"({tys},)" if len(args) == 1 else f"({tys})"
        rows.append(f"| `{name}` | `{res}` | `{tup}` |")
    rows += ["", TABLE_END]
    return "\n".join(rows)


def integration_with_table(text, table):
    i, j = text.find(TABLE_BEGIN), text.find(TABLE_END)
    if i < 0 or j < 0:
        raise ValueError("INTEGRATION.md has no ccall table markers")
    return text[:i] + table + text[j + len(TABLE_END):]


def main(argv):
    abi = load_abi()
    jl = render_abi_jl(abi)
    doc = integration_with_table(open(INTEGRATION).read(), render_table(abi))
    if "--check" in argv:
        stale = [p for p, want in ((ABI_JL, jl), (INTEGRATION, doc))
                 if not os.path.exists(p) or open(p).read() != want]
        for p in stale:
            print(f"stale: {p}")
        return 1 if stale else 0
    open(ABI_JL, "w").write(jl)
    open(INTEGRATION, "w").write(doc)
    print(f"wrote {ABI_JL} ({len(entries(abi))} entry points) and the table of {INTEGRATION}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
