"""The command-line driver's argument and input handling (no device needed:
every case here exits before a device context is created)."""
import subprocess

import pytest

from cli_helpers import CLI, EXACT_PLAN, plan_file, write

pytestmark = pytest.mark.skipif(not CLI.exists(), reason="synchro_amd not built")


def run(*args):
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=60)


def test_help():
    r = run("-h")
    assert r.returncode == 0 and "--continue" in r.stdout and "-f, --file" in r.stdout


def test_missing_required_options():
    assert run("-c", "x.json").returncode == 1
    assert run("-f", "x.txt").returncode == 1


def test_exclusive_log_levels_and_continue_needs_output(tmp_path):
    inp = write(tmp_path, "in.txt", "2 2\n1 1 0 0\n")
    cfg = plan_file(tmp_path, EXACT_PLAN)
    assert run("-f", inp, "-c", cfg, "-v", "-q").returncode == 1
    assert run("-f", inp, "-c", cfg, "--continue").returncode == 1


@pytest.mark.parametrize("text,msg", [("2 2\n1 1 0 5\n", "range [0, 1]"), ("2 2\n1 1 0\n", "Expected 4 integers"),
                                      ("0 3\n", "greater than 0")])
def test_bad_input_exits_with_3(tmp_path, text, msg):
    inp = write(tmp_path, "in.txt", text)
    r = run("-f", inp, "-c", plan_file(tmp_path, EXACT_PLAN))
    assert r.returncode == 3 and msg in r.stderr


def test_bad_config_exits_with_2(tmp_path):
    inp = write(tmp_path, "in.txt", "2 2\n1 1 0 0\n")
    cfg = write(tmp_path, "bad.json", "not json")
    assert run("-f", inp, "-c", cfg).returncode == 2


def test_continue_with_everything_done_needs_no_device(tmp_path):
    inp = write(tmp_path, "in.txt", "2 2\n1 1 0 0\n")
    out = write(tmp_path, "out.txt", "0: [1, 1] ((Brute, 0))\n")
    r = run("-f", inp, "-c", plan_file(tmp_path, EXACT_PLAN), "-o", out, "--continue")
    assert r.returncode == 0 and out.read_text() == "0: [1, 1] ((Brute, 0))\n"


def test_jobs_option_is_checked(tmp_path):
    inp = write(tmp_path, "in.txt", "2 2\n1 1 0 0\n")
    cfg = plan_file(tmp_path, EXACT_PLAN)
    for bad in ("0", "65", "x", "2x"):
        r = run("-f", inp, "-c", cfg, "-j", bad)
        assert r.returncode == 1 and "jobs" in r.stderr
    assert "-j, --jobs" in run("-h").stdout and "--reference-word" in run("-h").stdout
