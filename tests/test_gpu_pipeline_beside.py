"""The Beam's first pass beside Eppstein (pipeline.cpp): a context alone in
its process runs the Beam's first pass on a helper context while Eppstein
runs, and cuts it at Eppstein's bound through a stop cell (read by the
one-workgroup beam up to 64 states, by the host loop of the device-wide beam
above).  The plan's results must be those of the stages run one after
the other, which is what a process with two contexts alive does.

The lone-context side runs in a child process (one context there); this
process holds two contexts while it computes the other side.
"""
import json
import subprocess
import sys

import pytest

from conftest import ROOT, golden
from cli_helpers import EXACT_PLAN

CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import synchrowords_amd as sw
cases = json.load(open(sys.argv[2]))
ctx = sw.Context(0)
out = []
for c in cases:
    r = ctx.pipeline(c["n"], c["k"], c["delta"], plan=c["plan"])
    out.append([r.non_synchro, r.lower, r.upper, r.eppstein_bound, r.beam_bound, r.word,
                [name for name, _ in r.times_ms]])
ctx.close()
print(json.dumps(out))
"""


def _two_copies(m: int, seed: int) -> list[int]:
    import synchrowords_amd as sw
    d = [int(x) for x in sw.random_dfa(m, 2, seed)]
    return d + [x + m for x in d]


def _cases():
    import synchrowords_amd as sw
    from synchrowords_amd import gen
    plain = json.dumps(EXACT_PLAN)
    fixed = json.loads(plain)
    fixed["algorithms"][2]["config"] = {"presort": "indeg", "dynamic": False, "beam_size": "4"}
    capped = json.loads(plain)
    capped["algorithms"][2]["config"]["max_iter"] = "6"
    wide = json.loads(plain)  # a second, wider Beam pass after the first
    wide["algorithms"][2]["config"]["beam_exact_ratio"] = 50
    out = []
    for s in range(12):
        out.append((48, 2, [int(x) for x in sw.random_dfa(48, 2, s)], None))
    for s in range(4):
        out.append((64, 3, [int(x) for x in sw.random_dfa(64, 3, 500 + s)], plain))
        out.append((20 + 7 * s, 2, [int(x) for x in sw.random_dfa(20 + 7 * s, 2, 900 + s)], json.dumps(fixed)))
        out.append((30 + s, 2, [int(x) for x in sw.random_dfa(30 + s, 2, 950 + s)], json.dumps(capped)))
        out.append((40 + s, 2, [int(x) for x in sw.random_dfa(40 + s, 2, 970 + s)], json.dumps(wide)))
        out.append((40, 2, _two_copies(20, 700 + s), plain))  # not synchronizing
    out.append((24, 2, [int(x) for x in gen.cerny(24)], plain))
    out.append((16, 16, [int(x) for x in sw.random_dfa(16, 16, 31)], plain))
    out.append((40, 20, [int(x) for x in sw.random_dfa(40, 20, 32)], plain))  # past the one-workgroup beam's k
    for s in range(3):  # the device-wide beam
        out.append((100, 3, [int(x) for x in sw.random_dfa(100, 3, 40 + s)], plain))
        out.append((200, 2, [int(x) for x in sw.random_dfa(200, 2, s)], plain))
        out.append((90 + s, 2, [int(x) for x in sw.random_dfa(90 + s, 2, 60 + s)], json.dumps(capped)))
        out.append((128, 2, _two_copies(64, 80 + s), plain))  # not synchronizing
    out.append((70, 2, [int(x) for x in gen.cerny(70)], plain))
    for c in golden("random_misc.json"):
        out.append((c["n"], c["k"], c["delta"], plain))
    return [{"n": n, "k": k, "delta": d, "plan": p} for n, k, d, p in out]


@pytest.mark.gpu
def test_beam_beside_eppstein_matches_stages_in_order(ctx, tmp_path):
    import synchrowords_amd as sw
    cases = _cases()
    path = tmp_path / "cases.json"
    path.write_text(json.dumps(cases))
    print(f"[beside] {len(cases)} plans in a child process with one context", flush=True)
    child = subprocess.run([sys.executable, "-c", CHILD, str(ROOT), str(path)], capture_output=True, text=True,
                           timeout=240)
    assert child.returncode == 0, child.stderr[-2000:]
    beside = json.loads(child.stdout.strip().splitlines()[-1])
    other = sw.Context(0)  # two contexts alive here: the stages run one after the other
    try:
        for c, got in zip(cases, beside):
            r = ctx.pipeline(c["n"], c["k"], c["delta"], plan=c["plan"])
            want = [r.non_synchro, r.lower, r.upper, r.eppstein_bound, r.beam_bound, r.word,
                    [name for name, _ in r.times_ms]]
            assert got == want, (c["n"], c["k"], got[:5], want[:5])
            if not r.non_synchro and c["n"] > 15:
                assert r.word is not None and sw.replay(c["n"], c["k"], c["delta"], r.word) == 1
            if c["plan"] and '"max_iter": "6"' in c["plan"] and not r.non_synchro:
                # Six depths at most: a depth of at most 6, or the bound it started from.
                assert r.beam_bound <= 6 or r.beam_bound == r.eppstein_bound, (c["n"], r.beam_bound)
    finally:
        other.close()
    # And the reference's own results where it has them (configs/exact.json).
    for c in golden("random_misc.json"):
        i = next(j for j, x in enumerate(cases) if x["delta"] == c["delta"])
        assert bool(beside[i][0]) == bool(c["non_synchro"]), c["name"]
        if not c["non_synchro"]:
            assert (beside[i][1], beside[i][2]) == (c["lower"], c["upper"]), c["name"]
            if c["n"] > 15:
                assert (beside[i][3], beside[i][4]) == (c["eppstein"], c["beam"]), c["name"]
