"""The Julia binding (exactclosure.jl_amd/julia) checked without Julia, which
this image lacks: abi.jl and the ccall table of the integration guide are
what gen_abi.py writes from the header; every library call in
ExactClosureAMD.jl names an entry point of abi.jl with the C arity; the
reference names the scripts call (navierstokes.jl:10-38, burgers.jl:146-167)
are defined."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "exactclosure.jl_amd", "julia")
sys.path.insert(0, JL)
import gen_abi  # noqa: E402


def _split_args(s):
    """Top-level comma-separated arguments of a call's argument text."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _calls(text, start):
    """(argument text) of every call whose opening parenthesis is at start."""
    depth = 0
    for k in range(start, len(text)):
        if text[k] == "(":
            depth += 1
        elif text[k] == ")":
            depth -= 1
            if depth == 0:
                return text[start + 1:k]
    raise ValueError("unbalanced call")


def _module_text():
    txt = open(os.path.join(JL, "ExactClosureAMD.jl")).read()
    txt = re.sub(r'"""[\s\S]*?"""', '""', txt)          # docstrings
    txt = re.sub(r'"[^"\n]*"', '""', txt)                # one-line strings
    return "\n".join(line.split("#")[0] for line in txt.splitlines())


def test_generated_files_are_current():
    r = subprocess.run([sys.executable, os.path.join(JL, "gen_abi.py"), "--check"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_abi_jl_covers_the_header():
    jl = open(os.path.join(JL, "abi.jl")).read()
    defined = set(re.findall(r"(?m)^(ec_\w+)\(", jl))
    bound = set(re.findall(r"ccall\(\(:(ec_\w+), libexactclosure\)", jl))
    declared = set(gen_abi.header_params())
    assert defined == declared == bound


def test_every_library_call_has_the_c_arity():
    arity = {name: len(args) for name, _, args in gen_abi.entries()}
    txt = _module_text()
    seen = set()
    # call(f, args...) passes the context handle first
    for m in re.finditer(r"(?<!function )\bcall\(", txt):
        args = _split_args(_calls(txt, m.end() - 1))
        fn = args[0].removeprefix("EC.")
        assert fn in arity, f"call({fn}, ...) names no entry point"
        assert len(args) - 1 + 1 == arity[fn], f"{fn}: {len(args) - 1} arguments + ctx, C has {arity[fn]}"
        seen.add(fn)
    # direct calls of the generated functions
    for m in re.finditer(r"(?<![\w.])(?:EC\.)?(ec_\w+)\(", txt):
        fn = m.group(1)
        assert fn in arity, f"{fn} is not an entry point"
        args = _split_args(_calls(txt, m.end() - 1))
        assert len(args) == arity[fn], f"{fn}: {len(args)} arguments, C has {arity[fn]}"
        seen.add(fn)
    assert len(seen) >= 30, sorted(seen)


@pytest.mark.parametrize("name", [
    "poissonsolver", "project!", "randomfield", "propose_timestep", "step_forwardeuler!", "step_wray3!",
    "stresstensor", "tensordivergence!", "tensordivergence_nonsym!", "divergence!", "pressuregradient!",
    "tophat", "gaussian", "composekernel", "kernelproduct", "get_kernels", "applyfilter!", "apply_ΔHπ",
    "getsetup", "dnsaid_project", "compute_errors", "spectrum", "dns_aided_les", "run_dns_aided_les",
    "compute_fractions", "compute_dissipation", "solve_smagorinsky", "sfs_tensors_volume", "sfs_tensors_surface",
    "tensordissipation_staggered!", "sfs_dissipation"])
def test_reference_names_are_defined(name):
    txt = _module_text()
    pat = re.escape(name)
    assert re.search(rf"(?m)^\s*(?:function\s+{pat}\(|{pat}\([^)]*\)\s*=(?!=))", txt), name


def test_entry_points_match_python_binding():
    """abi.jl's types come from _abi.py; the C parameter counts from the
    header agree with the ctypes argument lists for every symbol."""
    abi = gen_abi.load_abi()
    hp = gen_abi.header_params()
    assert set(hp) == set(abi._SIGS)
    for name, params in hp.items():
        assert len(params) == len(abi._SIGS[name][1]), name
