#!/usr/bin/env bash
# Build the REFERENCE's own search (SynchroLib sources read in place from the
# reference tree, as oracle/build_ref.sh does) against this library's
# SynchroLib-side adapters, one pair of binaries per (N, K):
#
#   tests/integration/_bin/seams_<N>_<K>   the reference built for the GPU
#       (GPU=(true), THREADS=(1): its implicit-trie leaf tests and transition
#       applies go through its CUDA entry points), with the four .cu files
#       replaced by synchrowords_amd/integration/sw_seams.cpp;
#   tests/integration/_bin/plugin_<N>_<K>  the reference's plan loop with the
#       Exact stage served by ExactGpu (synchrowords_amd/integration/
#       run_exact_gpu.cpp in place of jit/jitmain.cpp); Brute, Eppstein and
#       Beam are the reference's own.
#
# Both link synchrowords_amd/libsynchrowords_amd.so and the test driver
# oracle/ref_driver.cpp (input and output in the reference's formats).  They
# prove the boundary: the reference's headers compile against the adapters,
# and on the GPU its search gives its own lengths (tests/test_gpu_integration.py).
# Test infrastructure; no reference file is copied or modified.
#
# With PLAN "reduce" only the plugin is built, as plugin_reduce_<N>_<K>, with
# the settings of configs/exact_reduce.json (Reduce before Exact; the test
# runs it in two stages through oracle/ref_driver.cpp --reduce-out and
# --reduce-in, one binary for each N, as the reference's app re-enters its
# JIT for the reduced automaton).
#
# usage: tests/integration/build_ref_gpu.sh N K [PLAN]
set -euo pipefail
N=${1:?N}
K=${2:?K}
PLAN=${3:-exact}
REF=${SYNCHRO_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
OUT="$HERE/_bin"
if [ ! -d "$REF/synchrolib" ]; then
  echo "reference tree not found at $REF" >&2
  exit 3
fi
LIB="$ROOT/synchrowords_amd/libsynchrowords_amd.so"
if [ ! -f "$LIB" ]; then
  echo "build the library first (python -m synchrowords_amd.build)" >&2
  exit 3
fi
mkdir -p "$OUT"

# The configs/exact.json plan's settings, as oracle/build_ref.sh gives them.
COMMON=(
  '-D$DEFINES$=' '-D$EXACT_DEF$=' '-D$EXACT_UNDEF$=' '-D$BEAM_DEF$=' '-D$BEAM_UNDEF$='
  '-D$EPPSTEIN_DEF$=' '-D$EPPSTEIN_UNDEF$=' '-D$BRUTE_DEF$=' '-D$BRUTE_UNDEF$='
  '-D$REDUCE_DEF$=' '-D$REDUCE_UNDEF$='
  "-DAUT_N=($N)" "-DAUT_K=($K)"
  '-DUPPER_BOUND=(1ULL * (AUT_N - 1) * (AUT_N - 1) + 1)'
  '-DGPU_MEMORY=(1024 * 2)' '-DTHREADS=(1)' -DJIT
  -DCOMPILE_BEAM -DCOMPILE_BRUTE -DCOMPILE_EPPSTEIN -DCOMPILE_EXACT
  '-DMAX_N=(15)'
  '-DTRANSITION_TABLES=(false)' '-DFIND_WORD=(false)'
  '-DBEAM_SIZE=(std::log2(AUT_N))' '-DDYNAMIC=(true)'
  '-DMIN_BEAM_SIZE=(AUT_N * std::log2(AUT_N))'
  '-DMAX_BEAM_SIZE=(AUT_N * AUT_N * std::log2(AUT_N))'
  '-DBEAM_EXACT_RATIO=(0.01)' '-DMAX_ITER=(-1)' -DPRESORT_INDEG
  '-DDFS_SHORTCUT=(true)' '-DDFS=(true)' '-DSTRICT_MEMORY_LIMIT=(false)'
  '-DMAX_MEMORY=(12 * 1024)' '-DDFS_MIN_LIST_SIZE=(8192)'
)
if [ "$PLAN" = reduce ]; then
  COMMON+=(-DCOMPILE_REDUCE '-DMIN_N=(80)' '-DLIST_SIZE_THRESHOLD=(AUT_N * AUT_K * 8)'
    '-DBFS_SMALL_LIST_SIZE=(AUT_N * AUT_K * AUT_K * 8)')
else
  COMMON+=('-DBFS_SMALL_LIST_SIZE=(AUT_N * AUT_K)')
fi
FLAGS=(-std=c++17 -O2 -DNDEBUG -pthread -w -I "$REF" -I "$REF/jit" -I "$ROOT/include"
  -I "$ROOT/synchrowords_amd/integration")
LINK=(-pthread -L "$ROOT/synchrowords_amd" -lsynchrowords_amd -Wl,-rpath,'$ORIGIN/../../../synchrowords_amd')
CORE=(
  "$REF/synchrolib/algorithm/exact/exact.cpp"
  "$REF/synchrolib/algorithm/exact/meet_in_the_middle.cpp"
  "$REF/synchrolib/algorithm/exact/dfs.cpp"
  "$REF/synchrolib/data_structures/preprocessed_transitions.cpp"
  "$REF/synchrolib/utils/distribution.cpp"
  "$REF/synchrolib/utils/connectivity.cpp"
  "$ROOT/oracle/ref_driver.cpp"
)

build() {  # name, GPU value, extra sources...
  local name=$1 gpu=$2
  shift 2
  local obj="$OUT/obj_${name}"
  mkdir -p "$obj"
  local pids=() objs=()
  for s in "${CORE[@]}" "$@"; do
    local o="$obj/$(basename "$s" .cpp).o"
    objs+=("$o")
    g++ "${FLAGS[@]}" "${COMMON[@]}" "-DGPU=($gpu)" -c "$s" -o "$o" &
    pids+=($!)
  done
  for p in "${pids[@]}"; do wait "$p"; done
  g++ "${objs[@]}" -o "$OUT/$name.tmp" "${LINK[@]}"
  mv "$OUT/$name.tmp" "$OUT/$name"
}

if [ "$PLAN" = reduce ]; then
  build "plugin_reduce_${N}_${K}" false "$ROOT/synchrowords_amd/integration/run_exact_gpu.cpp"
else
  build "seams_${N}_${K}" true "$REF/jit/jitmain.cpp" "$ROOT/synchrowords_amd/integration/sw_seams.cpp"
  build "plugin_${N}_${K}" false "$ROOT/synchrowords_amd/integration/run_exact_gpu.cpp"
fi
