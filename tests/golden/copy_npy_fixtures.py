"""Copies the reference's .npy test files (tests/npy/*.npy: counter-valued arrays of 110, 10x11 and
2x5x11 elements in every integer, float and complex type, format versions 1 and 2, C and Fortran
order, both byte orders) into tests/golden/npy/ as data fixtures. Run in a container that has the
reference sources at /root/reference."""
from __future__ import annotations

import shutil
import sys
from pathlib import Path

SRC = Path("/root/reference/tests/npy")
DST = Path(__file__).resolve().parent / "npy"


def main() -> int:
    if not SRC.is_dir():
        print(f"{SRC} is missing", file=sys.stderr)
        return 1
    DST.mkdir(exist_ok=True)
    n = 0
    for f in sorted(SRC.glob("*.npy")):
        shutil.copyfile(f, DST / f.name)
        n += 1
    print(f"copied {n} files into {DST}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
