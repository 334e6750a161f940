"""The SynchroLib-side adapters (synchrowords_amd/integration/) compile
against the reference's own headers, with the settings the reference's JIT
gives a GPU build, and the binaries built from them link this library.
Skipped where the reference tree is absent (the GPU box)."""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path(os.environ.get("SYNCHRO_REFERENCE", "/root/reference"))
DEFS = ["-D$DEFINES$=", "-D$EXACT_DEF$=", "-D$EXACT_UNDEF$=", "-D$BEAM_DEF$=", "-D$BEAM_UNDEF$=",
        "-D$EPPSTEIN_DEF$=", "-D$EPPSTEIN_UNDEF$=", "-D$BRUTE_DEF$=", "-D$BRUTE_UNDEF$=",
        "-D$REDUCE_DEF$=", "-D$REDUCE_UNDEF$=", "-DAUT_N=(130)", "-DAUT_K=(3)",
        "-DUPPER_BOUND=(1ULL * (AUT_N - 1) * (AUT_N - 1) + 1)", "-DGPU_MEMORY=(1024 * 2)", "-DTHREADS=(1)",
        "-DJIT", "-DCOMPILE_EXACT", "-DCOMPILE_BEAM", "-DCOMPILE_BRUTE", "-DCOMPILE_EPPSTEIN", "-DMAX_N=(15)",
        "-DTRANSITION_TABLES=(false)", "-DFIND_WORD=(false)", "-DBEAM_SIZE=(std::log2(AUT_N))",
        "-DDYNAMIC=(true)", "-DMIN_BEAM_SIZE=(AUT_N * std::log2(AUT_N))",
        "-DMAX_BEAM_SIZE=(AUT_N * AUT_N * std::log2(AUT_N))", "-DBEAM_EXACT_RATIO=(0.01)", "-DMAX_ITER=(-1)",
        "-DPRESORT_INDEG", "-DDFS_SHORTCUT=(true)", "-DDFS=(true)", "-DSTRICT_MEMORY_LIMIT=(false)",
        "-DMAX_MEMORY=(12 * 1024)", "-DDFS_MIN_LIST_SIZE=(8192)", "-DBFS_SMALL_LIST_SIZE=(AUT_N * AUT_K)"]

needs_ref = pytest.mark.skipif(not (REF / "synchrolib").is_dir(), reason="reference tree absent")


@needs_ref
@pytest.mark.parametrize("src,gpu", [("sw_seams.cpp", "true"), ("run_exact_gpu.cpp", "false")])
def test_adapters_compile_against_reference_headers(src, gpu):
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Werror=return-type", "-I", str(REF), "-I", str(REF / "jit"),
           "-I", str(ROOT / "include"), "-I", str(ROOT / "synchrowords_amd" / "integration"), *DEFS,
           f"-DGPU=({gpu})", str(ROOT / "synchrowords_amd" / "integration" / src)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]


@needs_ref
def test_integration_binaries_link_the_library():
    bins = sorted((ROOT / "tests" / "integration" / "_bin").glob("*_[0-9]*_[0-9]*"))
    bins = [b for b in bins if b.is_file()]
    if not bins:
        pytest.skip("integration binaries not built (run __graft_entry__.build())")
    for b in bins:
        r = subprocess.run(["ldd", str(b)], capture_output=True, text=True, timeout=60)
        assert "libsynchrowords_amd.so =>" in r.stdout and "not found" not in r.stdout, (b.name, r.stdout)
