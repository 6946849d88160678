"""Build libbsz_hip.so and the bsim_hip CLI in-tree for gfx950.

    python -m bitstreamzoo_amd.build [--jobs N] [--verbose]

HIP sources (*.hip) are compiled by hipcc for --offload-arch=gfx950 with
packed FP32 instructions disabled; host sources (*.cpp) by hipcc as C++.
Objects go to bitstreamzoo_amd/_build/, the library and the CLI next to this
file.  Always recompiles (a build is also the "does it build" check).
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libbsz_hip.so")
CLI = os.path.join(PKG, "bsim_hip")
ROOT = os.path.dirname(PKG)

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("BSZ_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", f"-I{os.path.join(ROOT, 'include')}"]
DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["frontend.cpp", "bsdata_io.cpp", "executor.cpp", "run_json.cpp", "capi.cpp",
               "fused.cpp", "kernels.hip"]
EXTRA_LIB_SOURCES_GLOB = ("fused_", "mlp_")  # more translation units picked up by prefix
CLI_SOURCES = ["bsim_hip.cpp"]


def _sources():
    srcs = list(LIB_SOURCES)
    for f in sorted(os.listdir(CSRC)):
        if f.startswith(EXTRA_LIB_SOURCES_GLOB) and f.endswith((".cpp", ".hip")) and f not in srcs:
            srcs.append(f)
    return srcs


def _compile(src: str, verbose: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
    cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip", *DEVICE_FLAGS]
    else:
        cmd[1:1] = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 0, verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    srcs = _sources() + CLI_SOURCES
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = dict(zip(srcs, ex.map(lambda s: _compile(s, verbose), srcs)))
    lib_objs = [objs[s] for s in srcs if s not in CLI_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *lib_objs,
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    tmp = CLI + ".tmp"
    cmd = [HIPCC, "-o", tmp, *[objs[s] for s in CLI_SOURCES], f"-L{PKG}", "-lbsz_hip",
           "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed (bsim_hip)\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CLI)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    build(a.jobs, a.verbose)
    print(f"built {LIB} and {CLI}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
