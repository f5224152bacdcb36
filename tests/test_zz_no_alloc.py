"""yee_get_scattered_field_noalloc (no device allocation: a fixed 32 MiB staging array per
device, grids in chunks of 4 Mi values) against yee_get_scattered_field, through the C ABI, bit
for bit: two calls in a row, the second on a smaller grid, then a grid larger than one chunk;
FP64 and FP32.
"""
import ctypes
import sys

import numpy as np
import pytest

from fdtd3d_amd.config import parse_args
from fdtd3d_amd.engine import COMP_INDEX, Engine, load_library

pytestmark = pytest.mark.gpu

COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(n, steps, precision):
    cfg = parse_args((f"--3d --size x:{n[0]},y:{n[1]},z:{n[2]} --use-pml --pml-size x:4 --same-size-pml --use-tfsf "
                      "--tfsf-size-left x:8 --tfsf-size-right x:8 --same-size-tfsf --angle-teta 30 --angle-phi 45 "
                      f"--angle-psi 90 --eps-sphere 3 --eps-sphere-center x:{n[0] // 2},y:{n[1] // 2},z:{n[2] // 2} "
                      f"--eps-sphere-radius 5 --stair-sphere-approx --time-steps {steps}").split())
    cfg.precision = precision
    eng = Engine(cfg)
    eng.set_material_sphere("eps", 3.0, 1.0, [n[0] // 2, n[1] // 2, n[2] // 2], 5, stair=True)
    eng.prepare()
    eng.step(steps)
    eng.sync()
    return eng


def _noalloc(eng, comp):
    out = np.full(eng.shape, np.nan)
    rc = load_library().yee_get_scattered_field_noalloc(eng.h, COMP_INDEX[comp], out.ctypes.data)
    assert rc == 0, load_library().yee_last_error()
    return out


@pytest.mark.parametrize("precision", [8, 4])
def test_noalloc_scattered_field_matches_allocating_entry_point(precision, engine_lib):
    # two calls in a row, the second on a smaller grid, then one larger than a 4 Mi-value chunk
    for n, steps in (((48, 40, 56), 40), ((30, 26, 34), 30), ((176, 160, 160), 25)):
        say(f"[gpu] no-alloc scattered field precision {precision} grid {n}")
        eng = _run(n, steps, precision)
        for comp in COMPS:
            ref = eng.scattered(comp)
            got = _noalloc(eng, comp)
            assert np.abs(ref).max() > 0, comp
            assert np.array_equal(got, ref), (n, comp, float(np.nanmax(np.abs(got - ref))))
        eng.close()


def test_noalloc_scattered_field_errors(engine_lib):
    eng = _run((24, 24, 24), 5, 8)
    L = load_library()
    assert L.yee_get_scattered_field_noalloc(eng.h, COMP_INDEX["Ex"], None) != 0
    assert L.yee_get_scattered_field_noalloc(eng.h, 9, np.zeros(1).ctypes.data) != 0
    eng.close()
    cfg = parse_args("--3d --size x:16 --same-size --time-steps 3".split())
    e2 = Engine(cfg)
    e2.prepare()
    out = np.zeros(cfg.shape)
    assert L.yee_get_scattered_field_noalloc(e2.h, COMP_INDEX["Ex"], out.ctypes.data) != 0   # no TF/SF
    assert b"TF/SF" in L.yee_last_error()
    e2.close()
