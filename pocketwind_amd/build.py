"""Compile libpocketwind.so (HIP, gfx950) in-tree with hipcc.

Run ``python -m pocketwind_amd.build`` or call :func:`build`. The library is written next to
this file so it travels with the source tree; it is not installed anywhere else.
"""
from __future__ import annotations

import concurrent.futures
import glob
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpocketwind.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 instructions are disabled for every kernel; the host pass ignores the feature.
COMPILE_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
]


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def _compile(src: str, obj: str) -> tuple[str, subprocess.CompletedProcess]:
    cmd = [HIPCC, *COMPILE_FLAGS, "-I", os.path.join(HERE, "..", "include"), "-c", "-o", obj, src]
    return src, subprocess.run(cmd, capture_output=True, text=True)


NAPI_DIR = os.path.join(HERE, "napi")
NAPI_ADDON = os.path.join(NAPI_DIR, "pw_napi.node")
NODE_INCLUDE = "/usr/include/node"


def build_napi(verbose: bool = False) -> str | None:
    """Compile the Node-API addon (plain C, gcc) against the built library, with an rpath to
    it; skipped, returning None, where the Node headers are absent."""
    if not os.path.isfile(os.path.join(NODE_INCLUDE, "node_api.h")):
        if verbose:
            print("node_api.h not found; Node-API addon not built", flush=True)
        return None
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror", "-shared", "-fPIC",
           "-DNAPI_VERSION=8", "-I", NODE_INCLUDE, "-I", os.path.join(HERE, "..", "include"),
           "-o", NAPI_ADDON + ".tmp", os.path.join(NAPI_DIR, "pw_napi.c"),
           "-L", HERE, "-lpocketwind", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("gcc failed to build pw_napi.node")
    os.replace(NAPI_ADDON + ".tmp", NAPI_ADDON)
    if verbose:
        print(f"built {NAPI_ADDON}", flush=True)
    return NAPI_ADDON


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every csrc/*.hip to an object in parallel, then link the shared library."""
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with tempfile.TemporaryDirectory(prefix="pw_build_") as tmp:
        objs = [os.path.join(tmp, os.path.basename(s) + ".o") for s in srcs]
        with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
            results = list(pool.map(_compile, srcs, objs))
        failed = [(src, r) for src, r in results if r.returncode != 0]
        for src, r in results:
            warnings = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
            if verbose and warnings:
                sys.stderr.write(f"{os.path.basename(src)}:\n" + "\n".join(warnings[:40]) + "\n")
        if failed:
            for src, r in failed:
                sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed to build libpocketwind.so")
        link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs]
        r = subprocess.run(link, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("linking libpocketwind.so failed")
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", flush=True)
    build_napi(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
