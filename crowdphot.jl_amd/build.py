"""Build libcrowdphot_hip.so in-tree for gfx950 with hipcc (no GPU needed).

    python crowdphot.jl_amd/build.py [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcrowdphot_hip.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["context.hip", "simultaneous.hip", "lm_batch.hip", "multipass.hip", "detection.hip"]
CPP_SOURCES = ["simulate.cpp", "server.cpp", "diagnostics.cpp", "aperture.cpp", "host_copy.cpp"]
# packed FP32 instructions are kept out of every HIP object
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]
CPP_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-pthread", "-march=x86-64-v3"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return r.stderr


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        objs.append(o)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", "-o", o, os.path.join(CSRC, s)])
    for s in CPP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        objs.append(o)
        cmds.append(["g++"] + CPP_FLAGS + ["-c", "-o", o, os.path.join(CSRC, s)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for err in ex.map(_run, cmds):
            if verbose and err:
                sys.stderr.write(err)
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
