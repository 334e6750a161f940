"""Write tests/golden/c2_count.json: the exact pair count of the subset-check
microbench, 2^20 forward x 2^20 backward sets over n = 256 states
(bench.make_lists, seed 20251203), by the oracle's all-pairs count over
every host thread (about two minutes on 8 cores).

The semantics are the reference's list-vs-list checker
(synchrolib/data_structures/cuda/subsets_checker_kernel.cu:8-33): the total
of cnt_i = #{j : small_i subset of large_j} over all i.  Besides the total the
fixture keeps the count of every block of 2^16 forward sets (so a sample can
be checked without the whole count), the first 4096 rows' count, and
checksums of both lists (so a test can see that the generator still makes
the same lists).

usage: python tests/golden/make_c2_count.py
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import bench  # noqa: E402
import oracle_lib as orc  # noqa: E402

COUNT = 1 << 20
N = 256
SEED = 20251203
BLOCK = 1 << 16


def checksum(a: np.ndarray) -> str:
    """xor and wrapping sum of every word, as hex."""
    x = np.bitwise_xor.reduce(a.ravel())
    s = int(a.ravel().sum(dtype=np.uint64))
    return f"{int(x):016x}:{s:016x}"


def main() -> None:
    small, large, planted = bench.make_lists(COUNT, N, 0.08, 0.75, seed=SEED)
    t0 = time.perf_counter()
    total, per_row = orc.subset_count_mt(small, large, per_row=True)
    dt = time.perf_counter() - t0
    blocks = [int(per_row[i:i + BLOCK].sum(dtype=np.uint64)) for i in range(0, COUNT, BLOCK)]
    assert sum(blocks) == total
    planted_ok = all(per_row[i] >= 1 for i, _ in planted)
    out = {
        "source": "oracle all-pairs count (oracle_subset_count_mt) over bench.make_lists(2^20, 256, 0.08, 0.75, "
                  f"seed={SEED}); semantics of subsets_checker_kernel.cu:8-33",
        "count": COUNT, "n": N, "p_small": 0.08, "p_large": 0.75, "seed": SEED,
        "pair_count": int(total),
        "first_4096_rows": int(per_row[:4096].sum(dtype=np.uint64)),
        "block_rows": BLOCK, "block_counts": blocks,
        "planted": [[int(i), int(j)] for i, j in planted], "planted_rows_counted": bool(planted_ok),
        "checksum_small": checksum(small), "checksum_large": checksum(large),
        "seconds": round(dt, 1),
    }
    (Path(__file__).resolve().parent / "c2_count.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"pair count {total} in {dt:.1f} s")


if __name__ == "__main__":
    main()
