"""Golden vectors of the configs/exact_reduce.json plan (Brute, Eppstein,
Beam, Reduce, Exact), written by the reference compiled from its own sources
(oracle/build_ref.sh N K 1 reduce).

When the reference's Reduce shrinks the automaton (synchrolib/algorithm/
reduce/reduce.hpp:31-103) its run returns early and its app builds the
plan again for the reduced N and continues (app/main.cpp:57-66).  This
script does the same with two binaries: the first run writes the pending
reduction (oracle/ref_driver.cpp --reduce-out), a binary built for the
reduced N continues it (--reduce-in) with the remaining Exact.

Cases: automata of at least 80 states (the plan's min_n) from the C3, C5,
random_misc and README fixtures.  Output: tests/golden/reduce_plan.json.

usage: python3 tests/golden/make_reduce_golden.py"""
import json
import subprocess
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
GOLDEN = ROOT / "tests" / "golden"
REF = ROOT / "oracle" / "_ref"
PLAN = "Brute,Eppstein,Beam,Reduce,Exact"


def binary(n: int, k: int) -> Path:
    b = REF / f"ref_reduce_{n}_{k}_1"
    if not b.exists():
        subprocess.run([str(ROOT / "oracle" / "build_ref.sh"), str(n), str(k), "1", "reduce"], check=True)
    return b


def fields(line: str) -> dict:
    return dict(kv.split("=") for kv in line.split("|", 1)[1].split())


def run_case(c: dict) -> dict:
    n, k = c["n"], c["k"]
    with tempfile.TemporaryDirectory() as d:
        inp = Path(d) / "in.txt"
        inp.write_text(f"{k} {n}\n" + " ".join(str(x) for x in c["delta"]) + "\n")
        red = Path(d) / "reduce.txt"
        out = subprocess.run([str(binary(n, k)), str(inp), "--plan", PLAN, "--reduce-out", str(red)],
                             capture_output=True, text=True, check=True).stdout
        f = fields(out.strip().splitlines()[-1])
        row = {"name": c["name"], "n": n, "k": k, "delta": c["delta"], "reduced_n": None}
        if f.get("reduce_pending") == "1":
            rn = int(f["reduced_n"])
            head = red.read_text().split("\n", 1)[0].split()
            row["reduced_n"] = rn
            row["reduce_bfs_steps"] = int(head[2])
            out2 = subprocess.run([str(binary(rn, k)), "/dev/null", "--plan", PLAN, "--reduce-in", str(red)],
                                  capture_output=True, text=True, check=True).stdout
            f = fields(out2.strip().splitlines()[-1])
        row.update(non_synchro=int(f["non_synchro"]), lower=int(f["lo"]), upper=int(f["hi"]))
        return row


def main():
    cases = []
    for name in ("c3_random200.json", "c5_random300.json", "random_misc.json", "readme.json"):
        cases += [c for c in json.loads((GOLDEN / name).read_text())["cases"] if c["n"] >= 80]
    rows = []
    for c in cases:
        rows.append(run_case(c))
        print(rows[-1]["name"], rows[-1]["n"], rows[-1]["reduced_n"], rows[-1]["lower"], rows[-1]["upper"], flush=True)
    (GOLDEN / "reduce_plan.json").write_text(json.dumps({
        "source": "reference compiled by oracle/build_ref.sh N K 1 reduce (configs/exact_reduce.json settings), "
                  "two-stage runs through oracle/ref_driver.cpp --reduce-out / --reduce-in",
        "plan": PLAN.split(","), "cases": rows}, indent=1))


if __name__ == "__main__":
    main()
