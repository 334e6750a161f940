"""Inputs and plans for the command-line driver tests."""
import json
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "synchrowords_amd" / "synchro_amd"

# The reference's configs/exact.json, configs/exact_reduce.json and
# configs/readme_config.json plans, as data.
EXACT_PLAN = {
    "upper_bound": "1ULL * (AUT_N - 1) * (AUT_N - 1) + 1", "threads": 1, "gpu": True,
    "algorithms": [
        {"name": "Brute", "config": {"max_n": "15"}},
        {"name": "Eppstein", "config": {"transition_tables": False, "find_word": False}},
        {"name": "Beam", "config": {"presort": "indeg", "dynamic": True, "min_beam_size": "AUT_N * std::log2(AUT_N)",
                                    "max_beam_size": "AUT_N * AUT_N * std::log2(AUT_N)", "beam_exact_ratio": 0.01}},
        {"name": "Exact", "config": {"dfs_min_list_size": "8192", "bfs_small_list_size": "AUT_N * AUT_K",
                                     "dfs": True, "dfs_shortcut": True, "max_memory_mb": "12 * 1024"}}]}
EXACT_REDUCE_PLAN = {
    "upper_bound": "1ULL * (AUT_N - 1) * (AUT_N - 1) + 1", "threads": 1, "gpu": False,
    "algorithms": [
        {"name": "Brute", "config": {"max_n": "15"}},
        {"name": "Eppstein", "config": {"transition_tables": False, "find_word": False}},
        {"name": "Beam", "config": {"presort": "indeg", "dynamic": True, "min_beam_size": "AUT_N * std::log2(AUT_N)",
                                    "max_beam_size": "AUT_N * AUT_N * std::log2(AUT_N)", "beam_exact_ratio": 0.01}},
        {"name": "Reduce", "config": {"min_n": "80", "list_size_threshold": "AUT_N * AUT_K * 8"}},
        {"name": "Exact", "config": {"dfs_min_list_size": "8192", "bfs_small_list_size": "AUT_N * AUT_K * AUT_K * 8",
                                     "dfs": True, "dfs_shortcut": True, "max_memory_mb": "12 * 1024"}}]}
README_PLAN = {
    "upper_bound": "1ULL * AUT_N * AUT_N * AUT_N / 6", "threads": 1, "gpu": False,
    "algorithms": [
        {"name": "Brute", "config": {"max_n": "20"}},
        {"name": "Eppstein", "config": {"transition_tables": False, "find_word": True}},
        {"name": "Beam", "config": {"presort": "indeg", "beam_size": "10"}},
        {"name": "Reduce", "config": {"min_n": "80", "list_size_threshold": "AUT_N * 16"}},
        {"name": "Exact", "config": {"dfs_min_list_size": "10000", "bfs_small_list_size": "AUT_N * 16",
                                     "dfs": True, "dfs_shortcut": True, "max_memory_mb": "2 * 1024"}}]}


def automata_text(cases) -> str:
    out = []
    for c in cases:
        out.append(f"{c['k']} {c['n']}")
        out.append(" ".join(str(int(x)) for x in c["delta"]))
    return "\n".join(out) + "\n"


def write(tmp: Path, name: str, text: str) -> Path:
    p = tmp / name
    p.write_text(text)
    return p


def plan_file(tmp: Path, plan: dict) -> Path:
    return write(tmp, "plan.json", json.dumps(plan))
