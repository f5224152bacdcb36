"""Edge cases of the engine against the oracle (needs the GPU): grids whose rows are longer
than one fused workgroup (N_z > 1024: in-place half steps over the whole grid), the smallest
grids the reference accepts, a grid that is PML everywhere (no fused box), the t = 0 rule
(no field update, the point source still fires), step counts split over several calls, and
a point source off the grid (ignored, as getFieldValueOrNullByAbsolutePos ignores it).

Bars as in the parity tests: FP64 bit-identical without PML and relative L2 <= 1e-12 with it;
FP32 relative L2 of the E and H vectors <= 1e-5 against the FP64 oracle."""
import sys

import numpy as np
import pytest

from fdtd3d_amd.config import parse_args
from fdtd3d_amd.engine import Engine, run_engine
from oracle.binding import run_oracle
from tests.helpers import check_fp32, field_names, rel_l2

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_FP32 = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(cfg, eng, o, exact):
    if cfg.precision == 4:
        # the E and H vectors, and each component holding >= 10 % of its vector (a component can
        # hold only FP64 rounding residue, ~1e-19)
        check_fp32("edge case", field_names(cfg), eng.field, o.field, TOL_FP32)
        return
    for f in field_names(cfg):
        got, want = eng.field(f), o.field(f)
        if exact:
            assert np.array_equal(got, want), (f, float(np.abs(got - want).max()))
        else:
            assert rel_l2(got, want) <= TOL_FP64, (f, rel_l2(got, want))


CASES = [
    # rows longer than 1024 cells: no fused kernel, in-place half steps; PML + oblique TF/SF
    ("long_rows", "--3d --size x:8,y:7,z:1100 --use-pml --pml-size x:1,y:1,z:10 --use-tfsf "
     "--tfsf-size-left x:3,y:3,z:20 --tfsf-size-right x:3,y:3,z:20 --angle-teta 20 --angle-phi 35 "
     "--angle-psi 60 --time-steps 40"),
    # N_z = 1025 with a point source and eps blocks
    ("rows_1025", "--3d --size x:6,y:6,z:1025 --point-source x:3,y:2,z:1000 --point-source-ex "
     "--point-source-hy --time-steps 30"),
    # the smallest grid: one updated cell
    ("min_3d", "--3d --size x:3,y:3,z:3 --point-source x:1,y:1,z:1 --point-source-ez --point-source-hx --time-steps 9"),
    ("min_tmz", "--2d-tmz --size x:3,y:3 --point-source x:1,y:1 --point-source-ez --time-steps 9"),
    ("small_uneven", "--3d --size x:4,y:3,z:5 --point-source x:2,y:1,z:3 --point-source-ey --time-steps 12"),
    # PML everywhere: no PML-free box for the fused kernel
    ("all_pml", "--3d --size x:8,y:8,z:8 --use-pml --pml-size x:4,y:4,z:4 --point-source x:4,y:3,z:4 "
     "--point-source-ez --time-steps 25"),
    ("all_pml_tmz", "--2d-tmz --size x:10,y:8 --use-pml --pml-size x:5,y:4 --point-source x:5,y:4 "
     "--point-source-ez --time-steps 25"),
    # TF/SF box one cell wide inside a thin PML-free box
    ("thin_tfsf", "--3d --size x:14,y:13,z:15 --use-pml --pml-size x:3,y:3,z:3 --use-tfsf "
     "--tfsf-size-left x:6,y:6,z:7 --tfsf-size-right x:7,y:6,z:7 --angle-teta 40 --angle-phi 20 "
     "--angle-psi 80 --time-steps 30"),
    # t = 0 performs no field update; the point source still fires
    ("one_step", "--3d --size x:10,y:9,z:11 --use-pml --pml-size x:2,y:2,z:2 --point-source x:5,y:4,z:5 "
     "--point-source-ez --point-source-hz --time-steps 1"),
    ("zero_steps", "--3d --size x:10,y:9,z:11 --use-tfsf --tfsf-size-left x:3,y:3,z:3 "
     "--tfsf-size-right x:3,y:3,z:3 --time-steps 0"),
    # TMz with a resonant Drude square inside PML + TF/SF (the single-run TMz half-step kernels, MODE 2)
    ("tmz_drude", "--2d-tmz --size x:40,y:36 --use-pml --pml-size x:6,y:6 --use-tfsf --tfsf-size-left x:9,y:8 "
     "--tfsf-size-right x:9,y:8 --angle-teta 90 --angle-phi 30 --angle-psi 90 --use-metamaterials --time-steps 300"),
    # TEz: the smallest grid, PML everywhere, a resonant Drude square with PML + TF/SF
    ("min_tez", "--2d-tez --size x:3,y:3 --point-source x:1,y:1 --point-source-hz --point-source-ex "
     "--angle-psi 0 --time-steps 9"),
    ("all_pml_tez", "--2d-tez --size x:10,y:8 --use-pml --pml-size x:5,y:4 --point-source x:5,y:4 "
     "--point-source-hz --angle-psi 0 --time-steps 25"),
    ("tez_drude", "--2d-tez --size x:40,y:36 --use-pml --pml-size x:6,y:6 --use-tfsf --tfsf-size-left x:9,y:8 "
     "--tfsf-size-right x:9,y:8 --angle-teta 90 --angle-phi 30 --angle-psi 0 --use-metamaterials --time-steps 300"),
    # TMx / TEy (the y-z and x-z planes, run on the 3D kernels): smallest grids, a Drude square
    ("min_tmx", "--2d-tmx --size x:4,y:3,z:3 --point-source x:2,y:1,z:1 --point-source-ex --point-source-hy "
     "--angle-phi 90 --angle-psi 0 --time-steps 9"),
    ("min_tey", "--2d-tey --size x:3,y:5,z:3 --point-source x:1,y:3,z:1 --point-source-hy --angle-phi 0 "
     "--time-steps 9"),
    ("tmy_drude", "--2d-tmy --size x:40,y:4,z:36 --use-pml --pml-size x:6,z:6 --use-tfsf --tfsf-size-left x:9,z:8 "
     "--tfsf-size-right x:9,z:8 --angle-teta 30 --angle-phi 0 --angle-psi 0 --use-metamaterials --time-steps 300"),
    # the six 1D lines (the 3D half-step kernels with two axes of extent 1; z of extent 1 for the
    # lines along x and y), at the angles each layout asserts
    ("line_exhy", "--1d-exhy --size z:70 --use-pml --pml-size z:8 --use-tfsf --tfsf-size-left z:15 "
     "--tfsf-size-right z:15 --angle-teta 0 --angle-phi 0 --angle-psi 90 --time-steps 90"),
    ("line_exhz", "--1d-exhz --size y:64 --use-pml --pml-size y:7 --use-tfsf --tfsf-size-left y:14 "
     "--tfsf-size-right y:16 --angle-teta 90 --angle-phi 90 --angle-psi 0 --point-source y:40 --point-source-hz "
     "--time-steps 90"),
    ("line_eyhx", "--1d-eyhx --size z:66 --use-tfsf --tfsf-size-left z:12 --tfsf-size-right z:12 --angle-teta 0 "
     "--angle-phi 90 --angle-psi 90 --current-source z:30 --current-source-jy --time-steps 80"),
    ("line_eyhz", "--1d-eyhz --size x:72 --use-pml --pml-size x:9 --use-tfsf --tfsf-size-left x:15 "
     "--tfsf-size-right x:15 --angle-teta 90 --angle-phi 0 --angle-psi 0 --time-steps 90"),
    ("line_ezhx", "--1d-ezhx --size y:60 --use-pml --pml-size y:6 --point-source y:25 --point-source-ez "
     "--point-source-hx --angle-teta 90 --angle-phi 90 --angle-psi 90 --time-steps 80"),
    ("line_ezhy", "--1d-ezhy --size x:64 --use-pml --pml-size x:8 --use-tfsf --tfsf-size-left x:16 "
     "--tfsf-size-right x:14 --angle-teta 90 --angle-phi 0 --angle-psi 90 --current-source x:32 "
     "--current-source-mz --time-steps 90"),
    # a resonant Drude layer on a line along x; the smallest line; a long line along z
    ("line_eyhz_drude", "--1d-eyhz --size x:120 --use-pml --pml-size x:10 --use-tfsf --tfsf-size-left x:20 "
     "--tfsf-size-right x:20 --angle-teta 90 --angle-phi 0 --angle-psi 0 --use-metamaterials --time-steps 300"),
    ("min_line", "--1d-ezhx --size y:3 --point-source y:1 --point-source-ez --angle-teta 90 --angle-phi 90 "
     "--angle-psi 90 --time-steps 9"),
    ("long_line", "--1d-exhy --size z:3000 --use-pml --pml-size z:40 --use-tfsf --tfsf-size-left z:100 "
     "--tfsf-size-right z:100 --angle-teta 0 --angle-phi 0 --angle-psi 90 --time-steps 400"),
    # a point source off the grid is ignored
    ("source_off_grid", "--3d --size x:12,y:10,z:14 --use-pml --pml-size x:2,y:2,z:2 --point-source x:30,y:4,z:5 "
     "--point-source-ez --time-steps 15"),
]


def _materials(name, cfg):
    if name in ("tmz_drude", "tez_drude", "tmy_drude"):
        w0 = 2 * np.pi * 299792458.0 / cfg.wavelength
        om = np.zeros(cfg.shape)
        ga = np.zeros(cfg.shape)
        om[15:25, 13:23] = 1.5 * w0
        ga[15:25, 13:23] = 0.01 * w0
        return {"omegape": om, "omegapm": om, "gammae": ga, "gammam": ga}
    if name == "line_eyhz_drude":
        w0 = 2 * np.pi * 299792458.0 / cfg.wavelength
        om = np.zeros(cfg.shape)
        ga = np.zeros(cfg.shape)
        om[70:85] = 1.5 * w0
        ga[70:85] = 0.01 * w0
        return {"omegape": om, "omegapm": om, "gammae": ga, "gammam": ga}
    if name != "rows_1025":
        return {}
    rng = np.random.default_rng(11)
    nb = [(s + 7) // 8 for s in cfg.shape]
    eps = np.kron(rng.uniform(1.0, 5.0, size=nb), np.ones((8, 8, 8)))[: cfg.shape[0], : cfg.shape[1], : cfg.shape[2]]
    return {"eps": eps}


@pytest.mark.parametrize("name,args", CASES)
@pytest.mark.parametrize("precision", [8, 4])
def test_edge_case_vs_oracle(name, args, precision, engine_lib):
    cfg = parse_args(args.split())
    mats = _materials(name, cfg)
    o = run_oracle(cfg, mats)
    cfg.precision = precision
    say(f"[gpu] edge {name} precision {precision}")
    eng = run_engine(cfg, mats)
    _check(cfg, eng, o, exact=not cfg.use_pml)
    if name == "zero_steps":
        for f in field_names(cfg):
            assert not np.any(eng.field(f)), f
    if name == "one_step":
        ez = eng.field("Ez")
        assert np.count_nonzero(ez) == 1 and ez[5, 4, 5] != 0
    eng.close()


@pytest.mark.parametrize("precision", [8, 4])
def test_steps_split_over_calls(precision, engine_lib):
    """yee_step (n) in several calls equals one call of the total (the step counter, the
    double-buffer parity and the incident grid continue across calls)."""
    cfg = parse_args("--3d --size x:40,y:36,z:44 --use-pml --pml-size x:5,y:4,z:6 --use-tfsf --tfsf-size-left x:9 "
                     "--tfsf-size-right x:8 --same-size-tfsf --angle-teta 30 --angle-phi 45 --angle-psi 90 "
                     "--point-source x:20,y:18,z:22 --point-source-hy --time-steps 41".split())
    cfg.precision = precision
    say(f"[gpu] split step calls precision {precision}")
    a = run_engine(cfg)
    b = Engine(cfg)
    b.prepare()
    for n in (1, 6, 13, 0, 21):
        b.step(n)
    b.sync()
    for f in field_names(cfg, aux=True):
        assert np.array_equal(a.field_raw(f), b.field_raw(f)), f
    ea, ha = a.incident(1000)
    eb, hb = b.incident(1000)
    assert np.array_equal(ea, eb) and np.array_equal(ha, hb)
    a.close()
    b.close()
