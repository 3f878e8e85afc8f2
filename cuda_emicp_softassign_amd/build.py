"""Builds libemsa.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Usage: python -m cuda_emicp_softassign_amd.build [--asm]
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libemsa.so")
SOURCES = ["capi.hip", "emicp.hip", "softassign.hip", "batched.hip", "adjoint.hip",
           "sa_unrolled.hip", "icp.hip", "probe.hip", "noalloc.hip", "emicp_cpu.cpp",
           "registration.cpp"]
ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# per source: the Softassign log-domain passes chain short matrix-core
# products into VALU work; their results go straight to VGPRs
SOURCE_FLAGS = {"softassign.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
                "emicp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-Wall",
          "-Wno-unused-function", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose: bool = False, host_flags=(), out: str = LIB) -> str:
    """Compiles every source and links `out` (default: the in-tree
    libemsa.so). host_flags are added to the host side of every compile and
    to the link (e.g. a sanitizer build: each flag preceded by -Xarch_host)."""
    objs = []
    objdir = os.path.join(HERE, "_obj") if out == LIB else out + "_obj"
    os.makedirs(objdir, exist_ok=True)
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        extra = SOURCE_FLAGS.get(os.path.basename(src), [])
        cmd = [HIPCC] + COMMON + DEVICE_FLAGS + extra + list(host_flags) + ["-c", src, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE_FLAGS + extra + list(host_flags) + \
                ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed for " + src)
        warn = "\n".join(l for l in r.stderr.splitlines()
                         if "packed-fp32-ops" not in l and l.strip())
        if warn and verbose:
            sys.stderr.write(warn + "\n")
        objs.append(obj)
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + \
        list(host_flags) + ["-Wl,-soname,libemsa.so"] + \
        ([] if host_flags else ["-Wl,--no-undefined"])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libemsa.so failed")
    os.replace(tmp, out)
    return out


TOOLS = os.path.join(HERE, "tools")
BIN = os.path.join(HERE, "bin")
# host programs: the registration harness (links libemsa) and the file tools
PROGRAMS = {"emsa_register": True, "add_rigid_transform": False, "merge": False,
            "pcd2ply": False}


def build_tools(verbose: bool = False) -> str:
    """Compiles the command-line harness and the file tools into bin/ with the
    host compiler (no device code)."""
    os.makedirs(BIN, exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    base = [cxx, "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + TOOLS]
    for prog, uses_lib in PROGRAMS.items():
        out = os.path.join(BIN, prog)
        cmd = base + [os.path.join(TOOLS, prog + ".cpp"), os.path.join(TOOLS, "cloud_io.cpp"),
                      "-o", out + ".tmp"]
        if uses_lib:
            cmd += ["-L" + HERE, "-lemsa", "-Wl,-rpath,$ORIGIN/..",
                    "-Wl,-rpath-link,/opt/rocm/lib", "-L/opt/rocm/lib", "-lamdhip64"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("build of %s failed" % prog)
        os.replace(out + ".tmp", out)
    return BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_tools(verbose="-v" in sys.argv))
