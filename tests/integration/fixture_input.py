"""Write the first COUNT automata of a golden fixture (tests/golden/NAME.json)
in the reference's input format ("K N" then the N*K transitions,
app/io.hpp:55-78), for the integration binaries.

usage: python3 tests/integration/fixture_input.py NAME COUNT > input.txt"""
import json
import sys
from pathlib import Path

GOLDEN = Path(__file__).resolve().parents[1] / "golden"


def lines(name: str, count: int):
    cases = json.loads((GOLDEN / f"{name}.json").read_text())["cases"][:count]
    for c in cases:
        yield f"{c['k']} {c['n']}"
        yield " ".join(str(x) for x in c["delta"])


if __name__ == "__main__":
    print("\n".join(lines(sys.argv[1], int(sys.argv[2]))))
