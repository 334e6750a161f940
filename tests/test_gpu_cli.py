"""The command-line driver on the device: the reference's README example
(data/readme_input.txt with configs/readme_config.json, expected lines
docs/install.md:101-105), the exact.json plan, and --continue."""
import re
import subprocess

import pytest

from cli_helpers import CLI, EXACT_PLAN, README_PLAN, automata_text, plan_file, write
from conftest import golden
from synchrowords_amd import replay

pytestmark = pytest.mark.gpu

LINE = re.compile(r"^(\d+): (?:NON SYNCHRO|\[(\d+), (\d+)\] \(((?:\(\w+, \d+\)(?:, )?)*)\)(?: \{([\d ]*)\})?)$")


def run(*args):
    r = subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r


def parse(text):
    out = []
    for line in text.splitlines():
        m = LINE.match(line)
        assert m, line
        idx, lo, hi, algos, word = m.groups()
        out.append((int(idx), None if lo is None else (int(lo), int(hi)),
                    re.findall(r"\((\w+), \d+\)", algos or ""), None if word is None else [int(x) for x in word.split()]))
    return out


@pytest.mark.parametrize("plan", [README_PLAN, EXACT_PLAN], ids=["readme_config", "exact"])
def test_readme_example(tmp_path, plan):
    cases = golden("readme.json")
    inp = write(tmp_path, "in.txt", automata_text(cases))
    out = tmp_path / "out.txt"
    run("-f", inp, "-c", plan_file(tmp_path, plan), "-o", out, "-q")
    lines = parse(out.read_text())
    assert [x[0] for x in lines] == list(range(len(cases)))
    names = [a["name"] for a in plan["algorithms"]]
    for (idx, bounds, algos, word), c in zip(lines, cases):
        if c["non_synchro"]:
            assert bounds is None
            continue
        assert bounds == (c["lower"], c["upper"]) == tuple(c["docs"])
        assert algos == names
        if word is not None:
            assert len(word) == bounds[1] and replay(c["n"], c["k"], c["delta"], word) == 1


def test_continue_appends_the_rest(tmp_path):
    cases = golden("c1_random48.json")[:6]
    inp = write(tmp_path, "in.txt", automata_text(cases))
    cfg = plan_file(tmp_path, EXACT_PLAN)
    full = tmp_path / "full.txt"
    run("-f", inp, "-c", cfg, "-o", full, "-q")
    part = tmp_path / "part.txt"
    part.write_text("".join(full.read_text().splitlines(keepends=True)[:2]))
    run("-f", inp, "-c", cfg, "-o", part, "--continue", "-q")
    a, b = parse(full.read_text()), parse(part.read_text())
    assert [(x[0], x[1]) for x in a] == [(x[0], x[1]) for x in b]
    for (idx, bounds, _, _), c in zip(a, cases):
        assert bounds == (c["lower"], c["upper"])


@pytest.mark.parametrize("jobs", [4, 8])
def test_jobs_keep_the_lines_in_order(tmp_path, jobs):
    """-j N solves N automata at a time (one context each) and still writes
    one line per automaton in input order, with the reference's bounds and
    words that replay; with --continue it resumes after the lines written."""
    cases = golden("c1_random48.json")[:16] + golden("readme.json")
    inp = write(tmp_path, "in.txt", automata_text(cases))
    cfg = plan_file(tmp_path, EXACT_PLAN)
    out = tmp_path / "out.txt"
    run("-f", inp, "-c", cfg, "-o", out, "-q", "-j", jobs)
    lines = parse(out.read_text())
    assert [x[0] for x in lines] == list(range(len(cases)))
    for (idx, bounds, algos, word), c in zip(lines, cases):
        if c["non_synchro"]:
            assert bounds is None
            continue
        assert bounds == (c["lower"], c["upper"])
        if word is not None:
            assert len(word) == bounds[1] and replay(c["n"], c["k"], c["delta"], word) == 1
    part = tmp_path / "part.txt"
    part.write_text("\n".join(out.read_text().splitlines()[:5]) + "\n")
    run("-f", inp, "-c", cfg, "-o", part, "-q", "-j", jobs, "--continue")
    assert [x[:2] for x in parse(part.read_text())] == [x[:2] for x in lines]


def test_reference_word_reproduces_the_readme_lines(tmp_path):
    """--reference-word: the word field is what the reference writes, so
    that with configs/readme_config.json every line of the README example
    equals docs/install.md:101-105 but for the times: the same bounds and
    algorithms, no word for the first three automata, and Eppstein's
    25-letter word for the fourth."""
    cases = golden("readme.json")
    inp = write(tmp_path, "in.txt", automata_text(cases))
    out = tmp_path / "out.txt"
    run("-f", inp, "-c", plan_file(tmp_path, README_PLAN), "-o", out, "-q", "--reference-word")
    lines = parse(out.read_text())
    names = [a["name"] for a in README_PLAN["algorithms"]]
    assert [x[0] for x in lines] == list(range(len(cases)))
    for (idx, bounds, algos, word), c in zip(lines, cases):
        if c["non_synchro"]:
            assert bounds is None
            continue
        assert bounds == tuple(c["docs"]) and algos == names
        if c.get("docs_eppstein_word"):
            assert word == c["docs_eppstein_word"]
        else:
            assert word is None
