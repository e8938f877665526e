"""Build recipe of the CPU oracle (test infrastructure only).

oracle/liboracle.so  the project's own C++ restatement (oracle.cpp)
oracle/_ref/libref_leaves.so  the reference's own leaf sources compiled where
    they lie under /root/reference/kernel/ocl with -DHOST: random.cc
    (LCG + samplers), hash.cc (sdbm), algebra/complex.cc, quaternion.cc. Only
    these compile as the reference ships them: any file that includes the
    transform headers needs the UNITTEST-only `Distance` template, and the
    UNITTEST build needs googletest, which the image lacks. A thin driver of
    our own (ref_leaves.cpp) exports them with C linkage. Built only when
    /root/reference exists.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_KERNEL = "/root/reference/kernel/ocl"
REF_SOURCES = ["random.cc", "hash.cc", "algebra/complex.cc", "algebra/quaternion.cc"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed: " + " ".join(cmd))


def build_oracle() -> str:
    out = os.path.join(HERE, "liboracle.so")
    tmp = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-Wall",
          os.path.join(HERE, "oracle.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_ref_leaves():
    if not os.path.isdir(REF_KERNEL):
        return None
    outdir = os.path.join(HERE, "_ref")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "libref_leaves.so")
    srcs = [os.path.join(REF_KERNEL, s) for s in REF_SOURCES] + [os.path.join(HERE, "ref_leaves.cpp")]
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-DHOST", "-I", REF_KERNEL, *srcs,
          "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build_oracle())
    print(build_ref_leaves())
