"""Write tests/golden/spice/<name>.ref.txt: what the reference's own SPICE
parser (netlist/spice_parser.cpp, compiled into oracle/_ref/libkaram_ref.so
by `make -C oracle ref`) makes of each tests/golden/spice/<name>.sp.  The
.sp texts are written for these tests; the .ref.txt files are the parser's
output (nodes, elements, model cards, directives) and are data only.

Run from the repository root:  python tests/golden/make_spice_golden.py
"""
from __future__ import annotations

import glob
import os
import sys

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle import ref as R  # noqa: E402

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spice")


def main() -> None:
    if not R.available():
        raise SystemExit("oracle/_ref/libkaram_ref.so is not built (make -C oracle ref)")
    for path in sorted(glob.glob(os.path.join(HERE, "*.sp"))):
        with open(path) as f:
            dump = R.spice_dump(f.read())
        out = path[:-3] + ".ref.txt"
        with open(out, "w") as f:
            f.write(dump)
        print(out, len(dump.splitlines()), "lines")


if __name__ == "__main__":
    main()
