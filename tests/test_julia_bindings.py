"""The Julia front-end (crowdphot.jl_amd/julia/CrowdPhotAMD.jl) against the C
ABI, without Julia: its mirrored structs have the header's field names, order
and C layout (the ctypes mirrors are checked against the header in
test_abi.py), and every library symbol it calls is declared in the header."""
import ctypes as C
import os
import re

from _support import load_pkg

pkg = load_pkg()
abi = pkg._abi

JL = os.path.join(os.path.dirname(os.path.abspath(pkg.__file__)), "julia", "CrowdPhotAMD.jl")

# Julia isbits types used in the mirrored structs: (size, alignment)
JL_TYPES = {"Int32": (4, 4), "Cint": (4, 4), "UInt32": (4, 4), "Int64": (8, 8), "UInt64": (8, 8),
            "UInt8": (1, 1), "Float32": (4, 4), "Float64": (8, 8)}

MIRRORS = {"CpPsfDesc": abi.PsfDesc, "CpSimOpts": abi.SimOpts, "CpSimResult": abi.SimResult,
           "CpLMOpts": abi.LMOpts, "CpMFResult": abi.MFResult, "CpTuning": abi.Tuning, "CpLMStat": abi.LMStat}


def _julia_structs():
    src = open(JL).read()
    out = {}
    for m in re.finditer(r"^(?:mutable )?struct (Cp\w+)\n(.*?)^end", src, re.S | re.M):
        body = "\n".join(line.split("#")[0] for line in m.group(2).splitlines())
        out[m.group(1)] = re.findall(r"(\w+)::([\w{}]+)", body)
    return src, out


def _layout(fields):
    """C offsets and size of a Julia isbits struct (Julia lays them out as C does)."""
    off, offs, align = 0, {}, 1
    for name, ty in fields:
        size, al = (8, 8) if ty.startswith("Ptr{") else JL_TYPES[ty]
        off = (off + al - 1) // al * al
        offs[name] = off
        off += size
        align = max(align, al)
    return offs, (off + align - 1) // align * align


def test_mirrored_structs_match_the_header():
    _, structs = _julia_structs()
    assert set(MIRRORS) <= set(structs), set(MIRRORS) - set(structs)
    for jname, py in MIRRORS.items():
        fields = structs[jname]
        names = [n for n, _ in fields]
        pnames = [n for n, _ in py._fields_ if not n.startswith("_")]
        assert names == pnames, (jname, names, pnames)
        offs, size = _layout(fields)
        for n in names:
            assert offs[n] == getattr(py, n).offset, (jname, n)
        assert size == C.sizeof(py), (jname, size, C.sizeof(py))


def test_called_symbols_are_declared():
    src, _ = _julia_structs()
    called = set(re.findall(r"sym\(:(cp_\w+)\)", src))
    assert len(called) >= 8, called
    assert called <= set(abi.SIGNATURES), called - set(abi.SIGNATURES)
