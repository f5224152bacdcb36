"""Near-to-far-field transform on the GPU (yee_ntff_snapshot / yee_ntff, fdtd3d-complex --use-ntff;
needs the GPU).

- The engine's diagrams against the oracle's restatement of Scheme::Pointing_scat over the
  same complex run (a dielectric sphere lit through TF/SF, PML around): forward diagram at the
  incident teta and off-plane elevations, several boxes.  FP64 within 1e-9 of the diagram's
  largest value (the device adds the face cells in another order and evaluates sin / cos with
  its own library); FP32 fields within 2e-3.
- The call contract: a real-field handle, a missing snapshot, a bad box are errors.
- The command line writes the reference's files (<name>_[timestep=t]_[step=s].txt, header,
  one line per angle) at every --interm-ntff-step, and the values equal the engine's.
"""
import dataclasses
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from fdtd3d_amd.config import parse_args
from fdtd3d_amd.engine import Engine, YeeError, ntff_angles, ntff_diagrams
from oracle.binding import oracle_ntff, run_oracle_complex, sphere_map

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fdtd3d_amd", "bin")

ARGS = ("--3d --size x:24 --same-size --use-pml --pml-size x:4 --same-size-pml --use-tfsf --tfsf-size-left x:7 "
        "--tfsf-size-right x:7 --same-size-tfsf --angle-teta 90 --angle-phi 0 --angle-psi 90 --wavelength 0.008 "
        "--dx 0.0005 --eps-sphere 2 --eps-sphere-center x:12,y:12,z:12 --eps-sphere-radius 3 --time-steps 120")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cfg(precision=8, extra=""):
    args = ARGS + extra
    cfg = dataclasses.replace(parse_args(args.split()), complex_fields=True)
    cfg.precision = precision
    return cfg


def _engine(cfg, steps):
    eng = Engine(cfg)
    eng.set_material_sphere("eps", cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius,
                            accuracy=cfg.sphere_accuracy, stair=cfg.stair_sphere)
    eng.prepare()
    eng.step(steps - 1)
    eng.ntff_snapshot()
    eng.step(1)
    eng.sync()
    return eng


BOXES = [([3, 3, 3], [22, 22, 22]), ([5, 4, 3], [19, 20, 21]), ([2, 2, 2], [22, 22, 22])]


@pytest.mark.parametrize("precision", [8, 4])
def test_ntff_vs_oracle(precision, engine_lib):
    cfg = _cfg(precision)
    say(f"[gpu] ntff vs oracle, precision {precision}")
    mats = {"eps": sphere_map(cfg, cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius)}
    re_, im_ = run_oracle_complex(dataclasses.replace(cfg, precision=8), mats)
    eng = _engine(cfg, cfg.time_steps)
    tol = 1e-9 if precision == 8 else 2e-3
    phis = ntff_angles(cfg)
    for left, right in BOXES:
        for teta in (math.pi / 2, 0.3, 2.2):
            got = eng.ntff(teta, phis, left, right)
            ref = np.array([oracle_ntff(re_, im_, teta, p, left, right)[0] for p in phis])
            scale = float(np.abs(ref).max())
            assert scale > 0
            err = float(np.abs(got - ref).max()) / scale
            say(f"[gpu]   box {left}..{right} teta {teta:.2f}: max {scale:.3e}, rel err {err:.2e}")
            assert err <= tol, (left, right, teta, err)
    eng.close()


def test_ntff_call_contract(engine_lib):
    cfg = _cfg()
    eng = Engine(cfg)
    eng.prepare()
    eng.step(3)
    with pytest.raises(YeeError):
        eng.ntff(math.pi / 2, [0.0], [3, 3, 3], [22, 22, 22])   # no snapshot one step before
    eng.ntff_snapshot()
    eng.step(2)
    with pytest.raises(YeeError):
        eng.ntff(math.pi / 2, [0.0], [3, 3, 3], [22, 22, 22])   # snapshot two steps before
    eng.ntff_snapshot()
    eng.step(1)
    assert eng.ntff(math.pi / 2, [0.0], [3, 3, 3], [22, 22, 22]).shape == (1,)
    for left, right in (([1, 3, 3], [22, 22, 22]), ([3, 3, 3], [23, 22, 22]), ([9, 3, 3], [9, 22, 22])):
        with pytest.raises(YeeError):
            eng.ntff(math.pi / 2, [0.0], left, right)
    eng.close()
    real = Engine(dataclasses.replace(cfg, complex_fields=False))
    real.prepare()
    with pytest.raises(YeeError):
        real.ntff_snapshot()
    real.close()
    odd = Engine(dataclasses.replace(parse_args("--3d --size x:23 --same-size --time-steps 4".split()),
                                     complex_fields=True))
    odd.prepare()
    with pytest.raises(YeeError):
        odd.ntff_snapshot()
    odd.close()


def _parse_ntff_file(path):
    with open(path) as fh:
        lines = fh.read().splitlines()
    vals = []
    for ln in lines[3:]:
        head, val = ln.split(", NTFF value = ")
        vals.append((head, float(head.split("angle NTFF = ")[1]), float(val)))
    return lines[:3], vals


def test_cli_ntff_files(tmp_path, engine_lib):
    exe = os.path.join(BIN, "fdtd3d-complex")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    args = (ARGS.replace("--time-steps 120", "--time-steps 130")
            + " --use-ntff --interm-ntff-step 60 --ntff-diff 2 --ntff-step-angle 30 --ntff-filename far").split()
    say("[gpu] cli ntff files")
    out = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stderr
    names = sorted(os.listdir(tmp_path))
    assert names == sorted(f"far_[timestep={t}]_[step={s}].txt" for t in (60, 120) for s in (0, 1))
    cfg = parse_args(args)
    cfg = dataclasses.replace(cfg, complex_fields=True)
    eng = _engine(cfg, 120)
    diagrams = ntff_diagrams(eng, cfg)
    eng.close()
    for s, left, right, phis, vals in diagrams:
        head, rows = _parse_ntff_file(os.path.join(tmp_path, f"far_[timestep=120]_[step={s}].txt"))
        assert head == [f"==== NTFF (step={s}) ====", "Forward diagram", ""]
        assert len(rows) == len(phis) == 13
        assert rows[0][0] == "timestep = 120, incident wave angle=(1.5708,0,1.5708,), angle NTFF = 0"
        got = np.array([r[2] for r in rows])
        assert np.allclose([r[1] for r in rows], phis, rtol=1e-5, atol=1e-6)
        assert np.allclose(got, vals, rtol=2e-5, atol=0)   # %g keeps 6 digits
    # real-field binaries refuse the option, as the reference's real build asserts
    bad = subprocess.run([os.path.join(BIN, "fdtd3d")] + args, capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert bad.returncode != 0 and "complex" in (bad.stdout + bad.stderr)


def test_cli_ntff_reverse_to_stdout(tmp_path, engine_lib):
    exe = os.path.join(BIN, "fdtd3d-complex")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    args = (ARGS.replace("--angle-phi 0", "--angle-phi 30")
            + " --use-ntff --interm-ntff-step 120 --ntff-reverse --ntff-to-stdout").split()
    say("[gpu] cli ntff reverse diagram")
    out = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    i = lines.index("==== NTFF (step=0) ====")
    assert lines[i + 1] == "Reverse diagram" and lines[i + 2] == ""
    assert lines[i + 3].startswith("timestep = 120, incident wave angle=(1.5708,0.523599,1.5708,), angle NTFF = 0.523599, ")
    assert not [n for n in os.listdir(tmp_path) if n.startswith("ntff-res")]
