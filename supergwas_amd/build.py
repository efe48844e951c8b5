"""In-tree build of the native parts (run by __graft_entry__.build()).

  lib/libkgwas_hip.so : gfx950 kernels + C ABI (hipcc)
  lib/libkgsynth.so   : host fill of the seeded synthetic bits (gcc)
  bin/kgwas           : drop-in `kgwas` executable (g++/hipcc link against libkgwas_hip)
  bin/kgwas_2d        : drop-in `kgwas_2d` all-pairs executable (same)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

ARCH = os.environ.get("KG_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed-FP32 VALU forms are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build_hip():
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    objs = []
    for src in ("kg_kernels.hip", "kg_api.cpp"):
        obj = os.path.join(BUILD, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-I", CSRC, "-I", INCLUDE, "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(LIB, "libkgwas_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_synth():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libkgsynth.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O3", "-fPIC", "-fopenmp", "-shared", "-std=c11", "-I", CSRC,
          os.path.join(CSRC, "synth_host.c"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_kgwas_bin():
    src = os.path.join(HOST, "kgwas_main.cpp")
    if not os.path.exists(src):
        return None
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, "kgwas")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-I", INCLUDE, "-I", HOST, src,
          "-L", LIB, "-lkgwas_hip", "-lz", "-Wl,-rpath,$ORIGIN/../lib", "-o", tmp])
    os.replace(tmp, out)
    src2 = os.path.join(HOST, "kgwas_2d_main.cpp")
    out2 = os.path.join(BIN, "kgwas_2d")
    _run(["g++", "-O3", "-std=c++17", "-I", INCLUDE, "-I", HOST, src2,
          "-L", LIB, "-lkgwas_hip", "-lz", "-Wl,-rpath,$ORIGIN/../lib", "-o", out2 + ".tmp"])
    os.replace(out2 + ".tmp", out2)
    return out


def build_all():
    build_synth()
    build_hip()
    build_kgwas_bin()


if __name__ == "__main__":
    build_all()
