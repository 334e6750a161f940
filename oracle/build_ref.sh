#!/usr/bin/env bash
# Compile the REFERENCE CPU path (SynchroLib, read in place from the reference
# tree) into oracle/_ref/ref_<N>_<K>_<T>, one binary per (N, K, threads), the
# same specialisation the reference's JIT makes (app/jit.hpp:25-56).
#
# Test infrastructure only: the binary is a checker and the CPU baseline,
# never part of the product path.
#
# The reference's JIT writes the plan's settings as #defines into its headers
# by replacing $PLACEHOLDER$ tokens (jitlib/jitlib.hpp:124-168,
# app/jit.hpp:86-141).  Here the same values are given as -D flags and every
# placeholder is defined empty; no reference file is copied or modified.
# The values below are the ones configs/exact.json produces:
#   upper_bound (N-1)^2+1; Brute max_n 15; Eppstein transition_tables false,
#   find_word false; Beam presort indeg, dynamic, min/max beam N*log2 N /
#   N*N*log2 N, beam_exact_ratio 0.01; Exact dfs, dfs_shortcut, 12 GiB,
#   dfs_min_list_size 8192, bfs_small_list_size N*K.
#
# With PLAN "reduce" the binary is ref_reduce_<N>_<K>_<T>, built with the
# settings of configs/exact_reduce.json: the same stages plus Reduce (min_n
# 80, list_size_threshold N*K*8) before Exact, whose bfs_small_list_size is
# N*K*K*8.
#
# usage: oracle/build_ref.sh N K [THREADS] [PLAN]
set -euo pipefail
N=${1:?N}
K=${2:?K}
T=${3:-1}
PLAN=${4:-exact}
REF=${SYNCHRO_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ "$PLAN" = reduce ]; then
  BIN="$OUT/ref_reduce_${N}_${K}_${T}"
else
  BIN="$OUT/ref_${N}_${K}_${T}"
fi
if [ ! -d "$REF/synchrolib" ]; then
  echo "reference tree not found at $REF" >&2
  exit 3
fi
mkdir -p "$OUT"
if [ -x "$BIN" ] && [ "$BIN" -nt "$HERE/ref_driver.cpp" ] && [ "$BIN" -nt "$0" ]; then
  exit 0
fi

DEFS=(
  '-D$DEFINES$=' '-D$EXACT_DEF$=' '-D$EXACT_UNDEF$=' '-D$BEAM_DEF$=' '-D$BEAM_UNDEF$='
  '-D$EPPSTEIN_DEF$=' '-D$EPPSTEIN_UNDEF$=' '-D$BRUTE_DEF$=' '-D$BRUTE_UNDEF$='
  '-D$REDUCE_DEF$=' '-D$REDUCE_UNDEF$='
  "-DAUT_N=($N)" "-DAUT_K=($K)"
  '-DUPPER_BOUND=(1ULL * (AUT_N - 1) * (AUT_N - 1) + 1)'
  '-DGPU=(false)' '-DGPU_MEMORY=(1024 * 2)' "-DTHREADS=($T)" -DJIT
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
  DEFS+=(-DCOMPILE_REDUCE '-DMIN_N=(80)' '-DLIST_SIZE_THRESHOLD=(AUT_N * AUT_K * 8)'
    '-DBFS_SMALL_LIST_SIZE=(AUT_N * AUT_K * AUT_K * 8)')
else
  DEFS+=('-DBFS_SMALL_LIST_SIZE=(AUT_N * AUT_K)')
fi
# x86-64-v3 (AVX2) instead of the reference's -march=native so the binary also
# runs on the host CPU of another machine.
FLAGS=(-std=c++17 -O3 -march=x86-64-v3 -DNDEBUG -pthread -w -I "$REF" -I "$REF/jit")
SRCS=(
  "$REF/jit/jitmain.cpp"
  "$REF/synchrolib/algorithm/exact/exact.cpp"
  "$REF/synchrolib/algorithm/exact/meet_in_the_middle.cpp"
  "$REF/synchrolib/algorithm/exact/dfs.cpp"
  "$REF/synchrolib/data_structures/preprocessed_transitions.cpp"
  "$REF/synchrolib/utils/distribution.cpp"
  "$REF/synchrolib/utils/connectivity.cpp"
  "$HERE/ref_driver.cpp"
)
OBJ="$OUT/obj_$(basename "$BIN")"
mkdir -p "$OBJ"
pids=()
objs=()
for s in "${SRCS[@]}"; do
  o="$OBJ/$(basename "$s" .cpp).o"
  objs+=("$o")
  g++ "${FLAGS[@]}" "${DEFS[@]}" -c "$s" -o "$o" &
  pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
g++ "${objs[@]}" -o "$BIN.tmp" -pthread
mv "$BIN.tmp" "$BIN"
rm -rf "$OBJ"
