"""Extracts the hand-written 1D golden vectors of the reference's unit test as a fixture.

Source (read as text, only its numbers are kept): Source/UnitTests/unit-test-internalscheme.cpp,
test1D_ExHy ... test1D_EzHy (:991-1600).  Each mode runs a 20-cell 1D grid (SIZE 20, TFSF_SIZE 7,
DX 0.002, LAMBDA 0.2, :33-38) with TF/SF on, 200 steps (`--time-steps 200 --use-tfsf`,
Tools/UnitTests/native/build-and-run-unit-test-internalscheme.sh), in a complex-field build whose
plane-wave source is sin(arg) + i cos(arg) (InternalScheme.inc.h:3309-3313).  For E_CENTERED the
test holds the complex E value of every cell after the run (:1033-1087 and the other modes).

Writes tests/golden/unit_test_1d_vectors.json:
    {mode: {"axis": "x"|"y"|"z", "component": "Ex"..., "angles": [teta, phi, psi],
            "re": [20 values], "im": [20 values]}}

    python tests/golden/extract_unit_test_vectors.py [/root/reference]
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
MODES = ["ExHy", "ExHz", "EyHx", "EyHz", "EzHx", "EzHy"]


def extract(ref_root: str) -> dict:
    path = os.path.join(ref_root, "Source", "UnitTests", "unit-test-internalscheme.cpp")
    with open(path) as fh:
        text = fh.read()
    out = {}
    for mode in MODES:
        start = text.index(f"void test1D_{mode} ()")
        end = text.index("#endif", text.index(f"test1D (intScheme.get", start))
        body = text[start:end]
        axis = re.search(r"CoordinateType ct1 = CoordinateType::([XYZ]);", body).group(1).lower()
        angles = [float(re.search(rf"FPValue angle{k} = ([-0-9.]+);", body).group(1)) for k in (1, 2, 3)]
        comp = re.search(r"test1D \(intScheme\.get(E[xyz]) \(\)", body).group(1)
        # the E_CENTERED block of expected values
        blk_start = body.index("layout_type == E_CENTERED")
        blk_end = body.index("layout_type == H_CENTERED")
        blk = body[blk_start:blk_end]
        vals = {}
        for m in re.finditer(r"FIELDVALUE \(([-0-9.e]+), ([-0-9.e]+)\), GRID_COORDINATE_1D \((\d+), ct1\)", blk):
            vals[int(m.group(3))] = (float(m.group(1)), float(m.group(2)))
        assert sorted(vals) == list(range(20)), (mode, sorted(vals))
        out[mode] = {"axis": axis, "component": comp, "angles": angles,
                     "re": [vals[i][0] for i in range(20)], "im": [vals[i][1] for i in range(20)]}
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = extract(ref)
    dst = os.path.join(HERE, "unit_test_1d_vectors.json")
    with open(dst, "w") as fh:
        json.dump(data, fh, indent=1)
    print(dst)


if __name__ == "__main__":
    main()
