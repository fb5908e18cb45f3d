"""Writes tests/golden/resample_lengths.json: the output lengths the reference's
own test expects from resample over input lengths 100..119 and up, down in 2..9
(impulse tests/test_signal.nim:149-256, `expected_result_lengths`), read as data
from that test file.  Run in the container that holds the reference tree:

    python tests/golden/make_signal_fixtures.py
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/tests/test_signal.nim"


def main():
    text = open(SRC).read()
    m = re.search(r"let expected_result_lengths = @\[(.*?)\]", text, re.S)
    if not m:
        raise SystemExit("expected_result_lengths not found")
    lengths = [int(v) for v in re.findall(r"\d+", m.group(1))]
    out = os.path.join(HERE, "resample_lengths.json")
    with open(out, "w") as f:
        json.dump({"source": "tests/test_signal.nim expected_result_lengths",
                   "signal": "arange(-50.0, 100.0)[:n]", "n": [100, 120], "up": [2, 10], "down": [2, 10],
                   "order": "n outer, then up, then down", "lengths": lengths}, f)
    print(out, len(lengths), "lengths")


if __name__ == "__main__":
    main()
