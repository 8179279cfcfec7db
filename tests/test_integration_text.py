"""The Julia ChaoslabHIPExt text in INTEGRATION.md against include/chaoslab_hip.h: its descriptor structs have the
header's field order and byte layout (the ctypes structures of chaos_amd/_lib.py, which test_abi_cpu.py ties to the
library), every system kind and reducer kind the library runs has a cl_system / cl_reducer method with the header's
number, and every ccall'ed symbol is declared in the header. Julia is not installed here, so the text is parsed."""
import ctypes
import os
import re

import pytest

from chaos_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL_SIZES = {"Int32": 4, "Int64": 8, "Float64": 8, "UInt8": 1, "Bool": 1}


def julia_text():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```julia\n(.*?)```", text, flags=re.S)
    ext = [b for b in blocks if "module ChaoslabHIPExt" in b]
    assert ext, "the ChaoslabHIPExt block is missing"
    return ext[0]


def header_defines():
    text = open(os.path.join(ROOT, "include", "chaoslab_hip.h")).read()
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"#define (CL_\w+) (\d+)", text)}


def jl_fields(src, name):
    m = re.search(r"struct %s\n(.*?)\nend" % name, src, flags=re.S)
    assert m, name
    out = []
    for decl in re.split(r";|\n", m.group(1)):
        decl = decl.strip()
        if "::" in decl:
            f, ty = decl.split("::", 1)
            out.append((f.strip(), ty.strip()))
    return out


def jl_size(ty, sizes):
    if ty in JL_SIZES:
        return JL_SIZES[ty], JL_SIZES[ty]
    if ty.startswith("Ptr{"):
        return 8, 8
    m = re.match(r"NTuple\{(\d+), (\w+)\}", ty)
    if m:
        n, (s, a) = int(m.group(1)), jl_size(m.group(2), sizes)
        return n * s, a
    if ty in sizes:
        return sizes[ty]
    raise AssertionError(f"unknown Julia type {ty}")


def jl_layout(src, name, sizes):
    """Offsets (C alignment, as Julia lays out isbits structs) and the size of a Julia struct of the text."""
    off, align, out = 0, 1, []
    for f, ty in jl_fields(src, name):
        s, a = jl_size(ty, sizes)
        off = (off + a - 1) // a * a
        out.append((f, off, s))
        off += s
        align = max(align, a)
    return out, (off + align - 1) // align * align


@pytest.mark.parametrize("jl,cs", [("CLSystem", L.cl_system), ("CLBirkhoff", L.cl_birkhoff),
                                   ("CLReducer", L.cl_reducer), ("CLChain", L.cl_chain), ("CLProbe", L.cl_probe_t)])
def test_descriptor_layouts(jl, cs):
    src = julia_text()
    sizes = {"CLReducer": (ctypes.sizeof(L.cl_reducer), 8), "CLBirkhoff": (ctypes.sizeof(L.cl_birkhoff), 8)}
    lay, size = jl_layout(src, jl, sizes)
    assert size == ctypes.sizeof(cs), (jl, size, ctypes.sizeof(cs))
    c_offsets = [getattr(cs, f).offset for f, *_ in cs._fields_]
    assert [o for _, o, _ in lay] == c_offsets, (jl, lay, c_offsets)


def test_every_kind_has_a_mapping():
    src = julia_text()
    d = header_defines()
    sys_kinds = {int(k) for k in re.findall(r"CLSystem\((\d+),", src)}
    red_kinds = {int(k) for k in re.findall(r"CLReducer\((\d+),", src)}
    # the systems of the propagate path (the per-orbit-K map has no reference type: it is this library's own)
    for name in ("CL_SYS_STANDARD_MAP", "CL_SYS_WRAPPED_STANDARD_MAP", "CL_SYS_STANDARD_MAP_CHAIN",
                 "CL_SYS_HENON_HEILES_GAUSS", "CL_SYS_PAIR_STANDARD_MAP", "CL_SYS_PAIR_WRAPPED_STANDARD_MAP"):
        assert d[name] in sys_kinds, name
    for name in ("CL_RED_QR", "CL_RED_BENETTIN", "CL_RED_MEGNO", "CL_RED_SALI", "CL_RED_FLI", "CL_RED_OFLI",
                 "CL_RED_GALI", "CL_RED_MONITOR", "CL_RED_RUNMONITOR", "CL_RED_STRETCHING", "CL_RED_BIRKHOFF",
                 "CL_RED_RLI", "CL_RED_PHASE_ADVANCE"):
        assert d[name] in red_kinds, name
    for alg in ("CGS", "MGS", "RGS", "HOUSEHOLDER", "AUGMENTED", "GIVENS"):
        assert re.search(r"qr_alg\(::%s\) = %d" % ({"HOUSEHOLDER": "HouseholderQR", "AUGMENTED": "AugmentedQR",
                                                     "GIVENS": "GivensQR"}.get(alg, alg), d["CL_QR_" + alg]), src)


def test_ccall_symbols_are_declared():
    src = julia_text()
    used = set(re.findall(r"sym\(:(cl_\w+)\)", src))
    assert {"cl_probe", "cl_create", "cl_propagate_chunk", "cl_state_bytes"} <= used
    assert used <= set(L.SIGNATURES), used - set(L.SIGNATURES)


def test_gpu_step_ns_reads_the_record():
    src = julia_text()
    assert "gpu_ns_per_step_wave" in src and "TOML.parsefile" in src
    assert not re.search(r"gpu_step_ns\([^)]*\)\s*=\s*[0-9.]+", src), "a constant in place of the measured record"
