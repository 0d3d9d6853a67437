"""The C ABI: library loads, exports every declared symbol, and the ctypes
structure layouts match the C header (CPU only; no compute calls)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from _support import ROOT, load_pkg

pkg = load_pkg()
abi = pkg._abi
HEADER = os.path.join(ROOT, "include", "crowdphot_hip.h")


def declared_symbols():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cp_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(syms) == set(abi.SIGNATURES), set(syms) ^ set(abi.SIGNATURES)


def test_struct_layouts_match_header():
    structs = {"cp_psf_desc": abi.PsfDesc, "cp_sim_opts": abi.SimOpts, "cp_sim_result": abi.SimResult,
               "cp_lm_opts": abi.LMOpts, "cp_lm_stat": abi.LMStat, "cp_request": abi.Request,
               "cp_request_stat": abi.RequestStat, "cp_tuning": abi.Tuning}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write("\n".join(lines))
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-o", exe, c], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.split() for l in out if l.strip())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)


def test_defaults_match_reference_keywords():
    lib = abi.load_library()
    o = abi.SimOpts()
    lib.cp_sim_opts_default(C.byref(o))
    ref = abi.sim_opts_default()
    for f, _ in abi.SimOpts._fields_:
        assert getattr(o, f) == getattr(ref, f), f
    lo = abi.LMOpts()
    lib.cp_lm_opts_default(C.byref(lo))
    lref = abi.lm_opts_default()
    for f, _ in abi.LMOpts._fields_:
        assert getattr(lo, f) == getattr(lref, f), f
    assert lib.cp_psf_nparams(abi.CP_PSF_CIRCULAR_GAUSSIAN) == 5
    assert lib.cp_psf_nparams(abi.CP_PSF_CIRCULAR_MOFFAT) == 6
    assert lib.cp_psf_nparams(abi.CP_PSF_GRIDDED_IMAGE) == 4


def test_request_bytes_estimate():
    lib = abi.load_library()
    b32 = lib.cp_request_bytes(4096, 4096, 200_000, 3.0, 12.0, abi.CP_PREC_MIXED, 1)
    b64 = lib.cp_request_bytes(4096, 4096, 200_000, 3.0, 12.0, abi.CP_PREC_FP64, 1)
    host = lib.cp_request_bytes(4096, 4096, 200_000, 3.0, 12.0, abi.CP_PREC_MIXED, 0)
    # images (6 float32 planes + mask), stamp rows and the stamp structure dominate
    assert 0.5e9 < b32 < 2e9 and b64 > b32 and host > b32
    assert lib.cp_request_bytes(0, 10, 5, 3.0, 12.0, 0, 0) == 0
    assert lib.cp_request_bytes(10, 10, 5, 0.0, 12.0, 0, 0) == 0


def test_server_argument_errors_without_gpu():
    lib = abi.load_library()
    h = C.c_void_p()
    assert lib.cp_server_create(C.byref(h), 0, 0, 0) == abi.CP_EARG
    assert lib.cp_server_create(None, 0, 0, 1) == abi.CP_EARG
    t = C.c_uint64()
    assert lib.cp_server_submit(None, None, C.byref(t)) == abi.CP_ESTATE
    assert lib.cp_server_poll(None, 1, None) == abi.CP_ESTATE


def test_null_context_is_rejected_without_gpu():
    lib = abi.load_library()
    assert lib.cp_synchronize(None) == abi.CP_ESTATE
    t = abi.Tuning()
    assert lib.cp_get_tuning(None, C.byref(t)) == abi.CP_ESTATE
    assert lib.cp_set_tuning(None, C.byref(t)) == abi.CP_ESTATE
    assert lib.cp_fit_simultaneous(None, None, None, None, 0, 0, None, None, None, 0, 1.0, None, None) == abi.CP_ESTATE
