"""The drop-in boundary, end to end: the reference's own search, compiled
from its sources with this library's SynchroLib-side adapters
(tests/integration/build_ref_gpu.sh), on the golden fixtures.

  seams_N_K   the reference built for the GPU (GPU=(true), THREADS=(1)): its
              implicit-trie leaf tests (subsets_implicit_trie.hpp:170-185)
              and transition applies (preprocessed_transitions.cpp:58-70)
              call its CUDA entry points, here sw_seams.cpp on sw_leaf_* and
              sw_transition_apply; the reference has no CPU path for them in
              this build, so its lengths come through the library;
  plugin_N_K  the reference's plan loop with ExactGpu as the Exact stage
              (run_exact_gpu.cpp): the reference's Brute, Eppstein and Beam,
              then sw_exact_solve.

Both must print the reference's [lower, upper] for every automaton."""
import os
import subprocess
import tempfile
from pathlib import Path

import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "integration" / "_bin"


def _run(binary: Path, cases, timeout: int):
    if not binary.exists():
        pytest.skip(f"{binary.name} not built (tests/integration/build_ref_gpu.sh needs the reference tree)")
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        for c in cases:
            n, k, d = c["n"], c["k"], c["delta"]
            f.write(f"{k} {n}\n" + " ".join(str(x) for x in d) + "\n")
        path = f.name
    try:
        out = subprocess.run([str(binary), path], capture_output=True, text=True, timeout=timeout,
                             env=dict(os.environ))
    finally:
        os.unlink(path)
    assert out.returncode == 0, out.stderr[-2000:]
    rows = []
    for line in out.stdout.splitlines():
        if "| non_synchro=" not in line:
            continue
        fields = dict(kv.split("=") for kv in line.split("|", 1)[1].split())
        rows.append((int(fields["non_synchro"]), int(fields["lo"]), int(fields["hi"])))
    assert len(rows) == len(cases), out.stdout[-2000:]
    return rows


def _check(rows, cases):
    for (ns, lo, hi), c in zip(rows, cases):
        assert ns == int(bool(c["non_synchro"])), c["name"]
        if not ns:
            assert (lo, hi) == (c["lower"], c["upper"]), (c["name"], lo, hi, c["lower"], c["upper"])


@pytest.mark.parametrize("fixture,n,count", [("c1_random48.json", 48, 12), ("c3_random200.json", 200, 2)])
def test_reference_search_through_the_seams(fixture, n, count):
    cases = [c for c in golden(fixture) if c["n"] == n][:count]
    _check(_run(BIN / f"seams_{n}_2", cases, timeout=240), cases)


@pytest.mark.parametrize("fixture,n,count", [("c1_random48.json", 48, 20), ("c3_random200.json", 200, 10),
                                              ("c5_random300.json", 300, 5)])
def test_reference_plan_with_the_exact_plugin(fixture, n, count):
    cases = [c for c in golden(fixture) if c["n"] == n][:count]
    _check(_run(BIN / f"plugin_{n}_2", cases, timeout=240), cases)


@pytest.mark.parametrize("name,n,reduced_n", [("random200_s0", 200, 166), ("random300_s1000", 300, 236)])
def test_reference_reduce_plan_with_the_exact_plugin(name, n, reduced_n):
    """configs/exact_reduce.json through the reference's own plan loop and
    Reduce, in two stages as its app runs them (app/main.cpp:57-66): the
    binary for N stops after the reference's Reduce with the reduced
    automaton and its list (--reduce-out); the binary for the reduced N
    continues (--reduce-in) with ExactGpu, which searches from that list
    (sw_exact_solve_from_list) and adds the Reduce's steps back.  Both the
    reduced N and the final bounds are the reference's
    (tests/golden/reduce_plan.json)."""
    case = [c for c in golden("reduce_plan.json") if c["name"] == name][0]
    assert case["reduced_n"] == reduced_n
    first, second = BIN / f"plugin_reduce_{n}_2", BIN / f"plugin_reduce_{reduced_n}_2"
    for b in (first, second):
        if not b.exists():
            pytest.skip(f"{b.name} not built (tests/integration/build_ref_gpu.sh N K reduce)")
    plan = "Brute,Eppstein,Beam,Reduce,Exact"
    with tempfile.TemporaryDirectory() as d:
        inp, red = Path(d) / "in.txt", Path(d) / "reduce.txt"
        inp.write_text(f"{case['k']} {n}\n" + " ".join(str(x) for x in case["delta"]) + "\n")
        out = subprocess.run([str(first), str(inp), "--plan", plan, "--reduce-out", str(red)], capture_output=True,
                             text=True, timeout=240)
        assert out.returncode == 0, out.stderr[-2000:]
        assert f"reduce_pending=1 reduced_n={reduced_n}" in out.stdout, out.stdout[-1000:]
        out = subprocess.run([str(second), "/dev/null", "--plan", plan, "--reduce-in", str(red)],
                             capture_output=True, text=True, timeout=240)
        assert out.returncode == 0, out.stderr[-2000:]
    line = [l for l in out.stdout.splitlines() if "| non_synchro=" in l][-1]
    fields = dict(kv.split("=") for kv in line.split("|", 1)[1].split())
    assert (int(fields["lo"]), int(fields["hi"])) == (case["lower"], case["upper"]), line
