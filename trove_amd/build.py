"""Builds libtrove_amd.so (HIP kernels for gfx950 + the C ABI) in-tree.

Each translation unit under trove_amd/csrc is compiled to an object in
trove_amd/_build/ (in parallel), then linked into trove_amd/libtrove_amd.so.
Run as `python -m trove_amd.build` or call build().
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD_DIR = os.path.join(PKG_DIR, "_build")
LIB_PATH = os.path.join(PKG_DIR, "libtrove_amd.so")

ARCH = os.environ.get("TROVE_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # Packed FP32 instructions stay off in every device compile.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE,
    "-I" + CSRC,
    "-Wall",
    "-Wno-unused-function",
]

HOST_FLAGS = ["-std=c++17", "-O3", "-fPIC", "-I" + INCLUDE, "-Wall"]


# Translation units instantiated per struct size are compiled once per size
# shard (-DTROVE_SHARD=k) so that the shards build in parallel.
SHARDED = {"copy.hip", "gather.hip"}
SIZE_SHARDS = 10  # TROVE_SIZE_SHARDS in csrc/common.h


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _compile(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines()
        if "packed-fp32-ops" not in line  # host pass does not know the device feature
    )
    return proc.returncode, " ".join(cmd), out


def build(jobs: int | None = None, verbose: bool = False) -> str:
    """Compiles every source and links the shared library; returns its path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    hip, cpp = _sources()
    cmds, objs = [], []
    for f in hip:
        shards = range(SIZE_SHARDS) if f in SHARDED else [None]
        for k in shards:
            suffix = "" if k is None else f".s{k}"
            obj = os.path.join(BUILD_DIR, f[:-4] + suffix + ".o")
            extra = [] if k is None else [f"-DTROVE_SHARD={k}"]
            cmds.append([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, f), "-o", obj])
            objs.append(obj)
    for f in cpp:
        obj = os.path.join(BUILD_DIR, f[:-4] + ".o")
        cmds.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, f), "-o", obj])
        objs.append(obj)
    jobs = jobs or min(len(cmds), max(1, min(16, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, cmds))
    failed = [r for r in results if r[0] != 0]
    for rc, cmd, out in results:
        if verbose or rc != 0:
            print(cmd)
            if out.strip():
                print(out)
    if failed:
        raise RuntimeError(f"{len(failed)} translation unit(s) failed to compile")
    tmp = LIB_PATH + ".tmp"
    link = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    rc, cmd, out = _compile(link)
    if rc != 0:
        print(cmd)
        print(out)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


DEVTESTS_SRC = [os.path.join(ROOT, "tests", "host", f) for f in ("devtests.hip", "queue_primitives.hip")]
DEVTESTS_LIB = os.path.join(ROOT, "tests", "host", "libtrove_devtests.so")


def build_devtests() -> str:
    """Compiles the device-API unit-test kernels (tests/host/devtests.hip)
    and the queue-primitive measurements (tests/host/queue_primitives.hip)
    into one test library."""
    tmp = DEVTESTS_LIB + ".tmp"
    rc, cmd, out = _compile([HIPCC, *HIP_FLAGS, "-shared", *DEVTESTS_SRC, "-o", tmp])
    if rc != 0:
        print(cmd)
        print(out)
        raise RuntimeError("device test library failed to compile")
    os.replace(tmp, DEVTESTS_LIB)
    return DEVTESTS_LIB


HOST_DRIVERS = ("trove_benchmark", "trove_block")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build_host_drivers() -> list:
    """Compiles the C++ host drivers (tests/host/trove_benchmark.cpp,
    trove_block.cpp): host code against the C ABI and the HIP runtime, linked
    to the in-tree libtrove_amd.so (call after build())."""
    out = []
    for name in HOST_DRIVERS:
        exe = os.path.join(ROOT, "tests", "host", name)
        tmp = exe + ".tmp"
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + INCLUDE,
               "-I" + os.path.join(ROCM, "include"), os.path.join(ROOT, "tests", "host", name + ".cpp"),
               "-L" + PKG_DIR, "-ltrove_amd", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
               "-Wl,-rpath,$ORIGIN/../../trove_amd", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", tmp]
        rc, line, text = _compile(cmd)
        if rc != 0:
            print(line)
            print(text)
            raise RuntimeError(f"{name} failed to compile")
        os.replace(tmp, exe)
        out.append(exe)
    return out


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
