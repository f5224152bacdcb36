"""The sphere-material entry point and the fdtd3d-compatible command line (need the GPU).

- yee_set_material_sphere must build the same eps grid as the reference's Scheme::initGrids:
  the fixture runs that used --eps-sphere are repeated with the sphere built by the engine,
  to the fixture parity bar.
- bin/fdtd3d runs a reference command line end to end and prints the "Total time" line.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from fdtd3d_amd.engine import Engine
from tests.helpers import field_names, golden_names, load_golden, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fdtd3d_amd", "bin")
TOL_FP64 = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


SPHERE_CASES = [n for n in golden_names() if "sphere" in n]


@pytest.mark.parametrize("name", SPHERE_CASES)
def test_sphere_material_matches_reference_fixture(name, engine_lib):
    cfg, _, data = load_golden(name)
    say(f"[gpu] sphere material {name}")
    eng = Engine(cfg)
    eng.set_material_sphere("eps", cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius,
                            accuracy=cfg.sphere_accuracy, stair=cfg.stair_sphere)
    eng.prepare()
    eng.step(cfg.time_steps)
    eng.sync()
    for f in field_names(cfg):
        ref = data[f].reshape(cfg.shape)
        r = rel_l2(eng.field(f), ref)
        assert r <= TOL_FP64, f"{name}/{f}: {r}"
    eng.close()


@pytest.mark.parametrize("binary", ["fdtd3d", "fdtd3d-f32"])
def test_cli_runs_reference_command_line(binary, engine_lib):
    exe = os.path.join(BIN, binary)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    args = ("--3d --size x:40 --same-size --use-pml --pml-size x:5 --same-size-pml --use-tfsf "
            "--tfsf-size-left x:10 --tfsf-size-right x:10 --same-size-tfsf --angle-teta 30 --angle-phi 45 "
            "--angle-psi 90 --eps-sphere 4 --eps-sphere-center x:20,y:20,z:20 --eps-sphere-radius 8 "
            "--sphere-accuracy 20 --time-steps 50 --save-res --log-level 2").split()
    say(f"[gpu] cli {binary}")
    out = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "Total time = " in out.stdout
    assert "Cell updates = 3136000 " in out.stdout   # 40^3 x 49

