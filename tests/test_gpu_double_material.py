"""--use-double-material-precision on the GPU: the engine against the oracle (needs the GPU).

The engine averages the fine eps / mu grids per component once, at yee_prepare, on the device
(8 fine cells in the reference's nested order), and its half-step kernels read the averages;
the PML tables come from the fine sigma profiles.  Bars as for the other parity tests: FP64
bit-identical without PML and within 1e-12 with it, FP32 within 1e-5 (vectors and every
component with >= 10 % of its vector's squared norm).  The material grids the engine holds
(fine spheres drawn on the host, random grids as uploaded) equal the oracle's bit for bit.
"""
import sys

import numpy as np
import pytest

from fdtd3d_amd.config import parse_args
from fdtd3d_amd.engine import Engine
from oracle.binding import Oracle, sphere_map
from tests.helpers import DOUBLE_MATERIAL_CASES, check_fp32, field_names, random_material, rel_l2

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _materials(cfg, files):
    mats = {name: random_material(cfg, 11 + s) for s, name in enumerate(files)}
    if cfg.eps_sphere != 1.0:
        mats["eps"] = sphere_map(cfg, cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius)
    if cfg.mu_sphere != 1.0:
        mats["mu"] = sphere_map(cfg, cfg.mu_sphere, 1.0, cfg.mu_sphere_center, cfg.mu_sphere_radius)
    return mats


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("args,files", DOUBLE_MATERIAL_CASES)
def test_double_material_precision_vs_oracle(args, files, precision, engine_lib):
    cfg = parse_args((args + " --use-double-material-precision").split())
    cfg.precision = precision
    say(f"[gpu] double material precision, precision {precision}: {args[:60]}")
    mats = _materials(cfg, files)
    o = Oracle(parse_args((args + " --use-double-material-precision").split()))
    for name, arr in mats.items():
        o.set_material(name, arr)
    o.prepare()
    o.step(cfg.time_steps)

    eng = Engine(cfg)
    assert eng.material_shape == cfg.material_shape
    for name, arr in mats.items():
        if name == "eps" and cfg.eps_sphere != 1.0:
            eng.set_material_sphere("eps", cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius,
                                    accuracy=cfg.sphere_accuracy, stair=cfg.stair_sphere)
        elif name == "mu" and cfg.mu_sphere != 1.0:
            eng.set_material_sphere("mu", cfg.mu_sphere, 1.0, cfg.mu_sphere_center, cfg.mu_sphere_radius,
                                    accuracy=cfg.sphere_accuracy, stair=cfg.stair_sphere)
        else:
            eng.set_material(name, arr)
    if precision == 8:
        # the fine grids as the engine holds them (host-drawn spheres included)
        for which, name in ((0, "eps"), (1, "mu")):
            np.testing.assert_array_equal(eng.material(name), o.material(which))
    eng.prepare()
    eng.step(cfg.time_steps)
    eng.sync()
    names = field_names(cfg)
    if precision == 8:
        for n in names:
            got, ref = eng.field(n), o.field(n)
            if cfg.use_pml:
                assert rel_l2(got, ref) <= 1e-12, (n, rel_l2(got, ref))
            else:
                assert np.array_equal(got, ref), (n, float(np.abs(got - ref).max()))
    else:
        check_fp32("double material " + args[:12], names, eng.field, o.field)
    assert max(float(np.abs(o.field(n)).max()) for n in names) > 0
    eng.close()
    o.close()


def test_double_material_precision_engine_rejects_what_the_reference_cannot_run(engine_lib):
    """Metamaterials with double material precision hit UNREACHABLE in the reference
    (Approximation.cpp:196): the engine refuses the configuration, and the adjoint and the
    batched TMz runs refuse the mode."""
    from fdtd3d_amd.config import OptionError
    from fdtd3d_amd.engine import YeeError
    base = "--3d --size x:12,y:12,z:12 --use-pml --pml-size x:3,y:3,z:3 --use-double-material-precision"
    with pytest.raises(OptionError):
        parse_args((base + " --use-metamaterials").split())
    cfg = parse_args(base.split())
    cfg.use_metamaterials = True
    with pytest.raises(YeeError):
        Engine(cfg)
