"""The 80^3 FP64 configuration in full on the GPU (accurate eps = 4 sphere of radius 20 at
(40, 40, 40), PML 10, TF/SF inset 20 at normal incidence, Courant 0.5, 200 steps):

  - the engine's sphere generator against the eps grid the reference wrote (bit for bit);
  - FP64: every component against the C oracle over the whole grid at 1e-12 relative L2, and
    against the reference's own summary (tests/golden/cfg1_sphere80.npz: L2, sum, largest
    magnitude, mid-z plane, 4096 seeded cells) at 1e-12;
  - FP32: every component within 1e-5 relative L2 of the FP64 oracle.
The oracle run takes about 80 s on one host core.
"""
import dataclasses
import sys

import numpy as np
import pytest

from fdtd3d_amd.engine import Engine
from oracle.binding import Oracle
from tests.helpers import load_summary, rel_l2, summary_errors

pytestmark = pytest.mark.gpu

COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(cfg, precision):
    eng = Engine(dataclasses.replace(cfg, precision=precision))
    eng.set_material_sphere("eps", cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius,
                            accuracy=cfg.sphere_accuracy, stair=False)
    eng.prepare()
    eng.step(cfg.time_steps)
    eng.sync()
    return eng


def test_cfg1_full_size_against_reference_and_oracle(engine_lib):
    cfg, data = load_summary("cfg1_sphere80")
    shape = cfg.shape
    say("[gpu] cfg1 80^3 FP64, 200 steps")
    eng = _engine(cfg, 8)
    eps = eng.material("eps")
    np.testing.assert_array_equal(eps.ravel(), data["Eps"])
    f64 = {c: eng.field(c) for c in COMPS}
    eng.close()
    errs = summary_errors(f64, data, shape)
    say(f"[gpu] cfg1 FP64 vs reference summary: { {c: max(e.values()) for c, e in errs.items()} }")
    for c, e in errs.items():
        for k, v in e.items():
            assert v <= 1e-12, (c, k, v)

    say("[gpu] cfg1 oracle, 200 steps (about 80 s)")
    o = Oracle(cfg)
    o.set_material("eps", eps)
    o.prepare()
    o.step(cfg.time_steps)
    ref = {c: o.field(c) for c in COMPS}
    for c in COMPS:
        r = rel_l2(f64[c], ref[c])
        say(f"[gpu] cfg1 FP64 {c} vs oracle: {r:.2e}")
        assert r <= 1e-12, (c, r)

    say("[gpu] cfg1 80^3 FP32, 200 steps")
    eng = _engine(cfg, 4)
    for c in COMPS:
        r = rel_l2(eng.field(c), ref[c])
        say(f"[gpu] cfg1 FP32 {c} vs FP64 oracle: {r:.2e}")
        assert r <= 1e-5, (c, r)
    eng.close()
