"""Build recipe of libhipmps (gfx950) and of the C++ visitor layer.

    python -m tnqvm_amd.build          # compiles everything in-tree

Products (git-ignored, shipped with the tree):
  tnqvm_amd/libhipmps.so       HIP kernels + engine + C ABI (include/hipmps.h)
  tnqvm_amd/libtnqvm_hipmps.so C++ TNQVM visitor layer (hip-mps) on top of libhipmps
  tests/cpp/bin/*              C++ visitor tests, and their ASan/UBSan builds (*_san)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhipmps.so")
VISLIB = os.path.join(PKG, "libtnqvm_hipmps.so")
ARCH = os.environ.get("HIPMPS_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed-FP32 instruction selection is disabled for all device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"),
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]

DEVICE_SOURCES = ["kernels.hip", "lq.hip", "kraus.hip", "engine.hip", "gates.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_hipmps(jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_visitor() -> str:
    """C++ hip-mps visitor layer (host code only, links libhipmps)."""
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cpp")] if os.path.isdir(HOST) else []
    if not srcs:
        return ""
    tmp = VISLIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + os.path.join(ROOT, "include"), "-I" + HOST,
          "-o", tmp] + srcs + ["-L" + PKG, "-lhipmps", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, VISLIB)
    return VISLIB


def build_cpp_tests() -> list:
    """The C++ visitor tests; they link the CPU oracle (oracle/liboracle.so) as their checker."""
    tdir = os.path.join(ROOT, "tests", "cpp")
    if not os.path.isdir(tdir):
        return []
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir])
    bdir = os.path.join(tdir, "bin")
    os.makedirs(bdir, exist_ok=True)
    out = []
    for f in sorted(os.listdir(tdir)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(bdir, os.path.splitext(f)[0])
        _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + HOST, "-o", exe,
              os.path.join(tdir, f), "-L" + PKG, "-ltnqvm_hipmps", "-lhipmps", "-L" + odir, "-loracle",
              "-Wl,-rpath," + PKG, "-Wl,-rpath," + odir])
        out.append(exe)
    return out


SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined", "-g", "-O1"]


def build_sanitized() -> list:
    """Host code under AddressSanitizer + UndefinedBehaviorSanitizer (g++, no device code): the
    visitor layer with the C++ visitor tests (their CPU cases), and the CPU oracle with a driver
    of its own (tests/cpp/sanitize/oracle_driver.cpp).  libhipmps itself is linked uninstrumented."""
    tdir = os.path.join(ROOT, "tests", "cpp")
    bdir = os.path.join(tdir, "bin")
    os.makedirs(bdir, exist_ok=True)
    odir = os.path.join(ROOT, "oracle")
    host = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cpp")]
    vis = os.path.join(bdir, "visitor_tests_san")
    _run(["g++", "-std=c++17"] + SAN_FLAGS + ["-I" + os.path.join(ROOT, "include"), "-I" + HOST, "-o", vis,
          os.path.join(tdir, "visitor_tests.cpp")] + host +
         ["-L" + PKG, "-lhipmps", "-L" + odir, "-loracle", "-Wl,-rpath," + PKG, "-Wl,-rpath," + odir])
    orc = os.path.join(bdir, "oracle_san")
    _run(["g++", "-std=c++17", "-fopenmp"] + SAN_FLAGS + ["-o", orc,
          os.path.join(tdir, "sanitize", "oracle_driver.cpp"), os.path.join(odir, "mps_oracle.cpp")])
    return [vis, orc]


def build_all(jobs: int = 4) -> None:
    build_hipmps(jobs)
    if build_visitor():
        build_cpp_tests()
        build_sanitized()


if __name__ == "__main__":
    build_all(int(sys.argv[1]) if len(sys.argv) > 1 else 4)
    print("built", LIB)
