"""Build recipe for libswm.so (HIP kernels for gfx950 + host C) and the CLIs.

Run ``python -m swmapper_amd.build`` (or ``__graft_entry__.build()``).  Every
call recompiles; outputs go to ``swmapper_amd/`` (the .so and the CLIs) and
``swmapper_amd/_build/`` (objects), both git-ignored.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libswm.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed-fp32 off for all HIP code (the host pass warns it does not know the
# feature; that is harmless)
HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-I" + INC, "-I" + CSRC,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-munsafe-fp-atomics",
]
C_FLAGS = ["-O3", "-fPIC", "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I" + INC, "-I" + CSRC]

HIP_SOURCES = ["swm_map.hip", "swm_index.hip", "swm_serve.hip"]
C_SOURCES = ["swm_io.c", "swm_synth.c", "swm_recipe.c", "swm_sam.c"]
CLIS = {"swmapper": "swm_cli_map.c", "swm_index": "swm_cli_index.c"}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        procs.append((subprocess.Popen([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o],
                                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True), src))
        objs.append(o)
    for src in C_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for p, src in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError("hipcc failed on " + src)
        if verbose and out:
            sys.stderr.write(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lz"])
    os.replace(tmp, LIB)
    for exe, src in CLIS.items():
        _run(["gcc", *C_FLAGS, "-o", os.path.join(HERE, exe), os.path.join(CSRC, src),
              "-L" + HERE, "-lswm", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
