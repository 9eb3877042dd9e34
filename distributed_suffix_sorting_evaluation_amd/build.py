"""Build recipe of the native parts (hipcc for gfx950, gcc for host-only code).

Outputs, in-tree next to this file:
  libdcx.so          HIP engine + C ABI of include/dcx.h (csrc/dcx_engine.hip)
  libdcx_textgen.so  seeded workload generators (csrc/textgen.c)
  dcx_cli            command-line driver mirroring lightweight_dcx's cli (csrc/cli.cpp)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DCX_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + INCLUDE]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build_engine():
    out = os.path.join(HERE, "libdcx.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "dcx_engine.hip")])
    os.replace(tmp, out)
    return out


def build_textgen():
    out = os.path.join(HERE, "libdcx_textgen.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O3", "-fPIC", "-fopenmp", "-shared", "-o", tmp, os.path.join(CSRC, "textgen.c"), "-lm"])
    os.replace(tmp, out)
    return out


def build_cli():
    src = os.path.join(CSRC, "cli.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(HERE, "dcx_cli")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-I" + INCLUDE, "-o", tmp, src,
          "-L" + HERE, "-ldcx", "-ldcx_textgen", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_all():
    build_textgen()
    build_engine()
    build_cli()


if __name__ == "__main__":
    build_all()
