"""Build libbfamx.so in-tree: host C with gcc, HIP kernels with hipcc for gfx950.

Usage: python -m bfam_amd.build [-j N]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbfamx.so")

ARCH = os.environ.get("BFAMX_OFFLOAD_ARCH", "gfx950")
INCS = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
CFLAGS = ["-O2", "-fPIC", "-std=gnu99", "-Wall", "-Wextra", "-Wno-unused-parameter"]
# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in generated code
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _sources():
    host = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host"))
                  if f.endswith(".c"))
    dev = sorted(os.path.join(CSRC, "device", f) for f in os.listdir(os.path.join(CSRC, "device"))
                 if f.endswith(".hip"))
    return host, dev


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l
                    and "(ignoring feature)" not in l)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    host, dev = _sources()
    cmds = []
    objs = []
    for src in host:
        o = os.path.join(OBJ, os.path.basename(src) + ".o")
        cmds.append(["gcc"] + CFLAGS + INCS + ["-c", src, "-o", o])
        objs.append(o)
    for src in dev:
        o = os.path.join(OBJ, os.path.basename(src) + ".o")
        cmds.append(["hipcc"] + HIPFLAGS + INCS + ["-c", src, "-o", o])
        objs.append(o)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run(["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lz", "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    args = ap.parse_args()
    print(build(args.j, args.v))
    sys.exit(0)
