"""Builds the CPU oracle (test infrastructure only) and, where the reference checkout is present,
the reference's own host code for cross-checks.

  oracle/_build/libppm_oracle.so  <- oracle.c (strict fp32, OpenMP)
  oracle/_ref/libref_host.so      <- mesh.cpp bvh.cpp bbox.cpp camera.cpp from the reference
                                     checkout, compiled where they lie, plus ref_driver.cpp

The reference's Makefile builds with -ffast-math; here it is built with strict IEEE float so
that its BVH and light table can be compared bit for bit with the product's.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")
REF_OUT = os.path.join(HERE, "_ref")
ORACLE_LIB = os.path.join(OUT, "libppm_oracle.so")
REF_LIB = os.path.join(REF_OUT, "libref_host.so")
REF_ROOT = os.environ.get("PPM_REFERENCE_ROOT", "/root/reference")
REF_SOURCES = ("mesh.cpp", "bvh.cpp", "bbox.cpp", "camera.cpp")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)


def build_oracle():
    os.makedirs(OUT, exist_ok=True)
    tmp = ORACLE_LIB + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
          "-Wall", os.path.join(HERE, "oracle.c"), "-o", tmp, "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_ref():
    """Returns the library path, or None when the reference checkout is absent."""
    srcs = [os.path.join(REF_ROOT, s) for s in REF_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    os.makedirs(REF_OUT, exist_ok=True)
    tmp = REF_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++11", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-w",
          "-I" + REF_ROOT] + srcs + [os.path.join(HERE, "ref_driver.cpp"), "-o", tmp])
    os.replace(tmp, REF_LIB)
    return REF_LIB


def build():
    build_oracle()
    return build_ref()


if __name__ == "__main__":
    print(build_oracle())
    print(build_ref())
