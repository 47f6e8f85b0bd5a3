"""Build libcompas_amd.so (HIP, gfx950) in-tree with hipcc.

    python compas-toolkit_amd/build.py            # full rebuild
    python compas-toolkit_amd/build.py --incremental

Every .hip/.cpp file under csrc/src is compiled for gfx950 and linked into
compas-toolkit_amd/libcompas_amd.so.  HIP code is compiled with packed FP32
instructions disabled (-packed-fp32-ops).
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "src")
INCLUDES = [os.path.join(HERE, "csrc", "include"), os.path.join(ROOT, "include")]
OBJ = os.path.join(HERE, "build", "obj")
LIB = os.path.join(HERE, "libcompas_amd.so")
ARCH = os.environ.get("COMPAS_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
HIP_ONLY = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def sources():
    out = []
    for name in sorted(os.listdir(SRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(SRC, name))
    return out


def headers_digest():
    h = hashlib.sha256()
    for inc in INCLUDES:
        for dirpath, _, files in os.walk(inc):
            for f in sorted(files):
                if f.endswith((".h", ".hpp")):
                    with open(os.path.join(dirpath, f), "rb") as fh:
                        h.update(fh.read())
    return h.hexdigest()


def compile_one(src, incremental, hdr):
    base = os.path.basename(src)
    obj = os.path.join(OBJ, base + ".o")
    stamp = obj + ".stamp"
    with open(src, "rb") as fh:
        key = hashlib.sha256(fh.read() + hdr.encode() + " ".join(COMMON + HIP_ONLY).encode()).hexdigest()
    if incremental and os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == key:
        return obj, None
    # hipcc compiles .cpp as HIP too: give every file the gfx950 target
    cmd = [HIPCC] + COMMON + [f"-I{i}" for i in INCLUDES] + HIP_ONLY
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f"compile failed: {base}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    with open(stamp, "w") as fh:
        fh.write(key)
    return obj, None


def build(incremental: bool = False, jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hdr = headers_digest()
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, min(16, os.cpu_count() or 4)))
    errors = []
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda s: compile_one(s, incremental, hdr), srcs):
            objs.append(obj)
            if err:
                errors.append(err)
    if errors:
        raise RuntimeError("\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    args = ap.parse_args()
    try:
        build(args.incremental, args.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
