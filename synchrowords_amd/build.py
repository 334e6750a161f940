"""Build the HIP library in-tree: ``python -m synchrowords_amd.build``.

Kernels (``csrc/*.hip``) are compiled by hipcc for gfx950; host sources
(``csrc/*.cpp``) by g++ against the HIP runtime; the result is
``synchrowords_amd/libsynchrowords_amd.so``, and the command-line driver
``synchrowords_amd/synchro_amd`` (``app/synchro_amd.cpp``) linked against it.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libsynchrowords_amd.so"
CLI = HERE / "synchro_amd"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("SW_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # No packed FP32 instructions in any kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-pass-failed",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-Wall",
             "-Wno-unused-parameter", "-Wno-unused-result"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build(verbose: bool = False) -> Path:
    hipcc = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    jobs = []
    for src in sorted(CSRC.glob("*.hip")):
        obj = objdir / (src.stem + ".o")
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in sorted(CSRC.glob("*.cpp")):
        obj = objdir / (src.stem + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(_run, jobs))
    objs = [str(objdir / (Path(j[-3]).stem + ".o")) for j in jobs]
    tmp = OUT.with_suffix(".so.tmp")
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *objs, f"-L{ROCM}/lib", "-lamdhip64"])
    os.replace(tmp, OUT)
    # The command-line driver, linked against the library beside it.
    cli_tmp = CLI.with_suffix(".tmp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{HERE.parent / 'include'}", str(HERE / "app" / "synchro_amd.cpp"),
          "-o", str(cli_tmp), f"-L{HERE}", "-lsynchrowords_amd", "-Wl,-rpath,$ORIGIN", f"-L{ROCM}/lib",
          "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(cli_tmp, CLI)
    if verbose:
        print(f"built {OUT} and {CLI}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
