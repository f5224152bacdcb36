"""Analytic test problems on the GPU (yee_set_test_functions / yee_get_diff_norm; needs the GPU).

- The engine against the oracle on the same cases as the oracle's bit-exact check against the
  reference driver: FP64 fields bit-identical (the callbacks are evaluated with the reference's
  operations), sin1 within 1e-12 (the device sin may differ from the C library's in the last
  bit); per-step norms and percentages within 1e-9 relative (the device adds the squares in
  another order).  FP32 only for sin1: the polynomial problems cancel the curl against the right
  side to ~1e-9 relative (2 c^2 x t against itself, leaving x^2), which FP32 (6e-8) cannot
  resolve; the reference's suite runs them in its default FP64 build.
- The reference suite's pass criteria (Tests/suite t2-t5) on the engine, FP64 and FP32.
- The command line prints the reference's "DIFF NORM" lines, which the suite's awk reads.
- The exp solutions of the 1D pairs (Tests/suite t7.x): engine against the oracle, real and
  complex handles; the suite's bars at full size on every pair; the complex build's line.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from fdtd3d_amd.config import parse_args
from fdtd3d_amd.engine import Engine
from oracle.binding import Oracle
from tests.helpers import EXP_CASES, TEST_FN_CASES, check_fp32, field_names, rel_l2, run_norms

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fdtd3d_amd", "bin")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("args,precision", [(a, 8) for a in TEST_FN_CASES] + [(a, 4) for a in TEST_FN_CASES if "sin1" in a])
def test_test_functions_vs_oracle(args, precision, engine_lib):
    cfg = parse_args(args.split())
    cfg.precision = precision
    say(f"[gpu] test functions precision {precision}: {args[:70]}")
    o = Oracle(parse_args(args.split()))
    o.prepare()
    ref = run_norms(o, cfg)
    eng = Engine(cfg)
    eng.prepare()
    got = run_norms(eng, cfg)
    names = field_names(cfg)
    assert sorted(got) == sorted(ref)
    if precision == 8:
        sin = "sin1" in args
        for n in names:
            if sin:
                assert rel_l2(eng.field(n), o.field(n)) <= 1e-12, n
            else:
                assert np.array_equal(eng.field(n), o.field(n)), (n, float(np.abs(eng.field(n) - o.field(n)).max()))
        for n in ref:
            # norm, percent, sum, max |exact| of every step
            assert np.allclose(got[n], ref[n], rtol=1e-9, atol=0), n
    else:
        check_fp32("test functions " + args[:10], names, eng.field, o.field)
    eng.close()
    o.close()


_FN = "--use-{k}-border-condition --use-{k}-start-values --use-{k}-right-side --calc-{k}-diff-norm"


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("precision", [8])
@pytest.mark.parametrize("scheme,kind,comps", [
    ("--3d", "polinom1", ["Ez"]), ("--2d-tmz", "polinom1", ["Ez"]), ("--2d-tey", "polinom1", ["Ez"]),
    ("--3d", "polinom2", ["Ex", "Ey", "Ez"]), ("--3d", "polinom3", ["Ez"]), ("--2d-tmz", "polinom3", ["Ez"]),
    ("--2d-tey", "polinom3", ["Ez"]),
])
def test_suite_polynomial_bar(scheme, kind, comps, precision, layout, engine_lib):
    """Tests/suite t2.x / t3 / t4.x: the largest diff-norm percentage over the run <= 1e-4 %, in
    both Yee layouts (t2.3 runs --layout-type 0 and 1)."""
    worst = 0.0
    for size, dx, steps in ((20, "0.001", 101), (40, "0.0005", 201)):
        args = (f"{scheme} --size x:{size} --same-size --angle-phi 0 --dx {dx} --wavelength 0.02 "
                + _FN.format(k=kind) + f" --time-steps {steps}")
        cfg = parse_args(args.split())
        cfg.precision = precision
        cfg.layout = layout
        eng = Engine(cfg)
        eng.prepare()
        norms = run_norms(eng, cfg)
        eng.close()
        worst = max(worst, max(float(norms[c][:, 1].max()) for c in comps))
    say(f"[gpu] suite {scheme} {kind} precision {precision} layout {layout}: largest percent {worst:.3e}")
    assert worst <= 1e-4


@pytest.mark.parametrize("precision", [8])
def test_suite_sin1_convergence(precision, engine_lib):
    """Tests/suite t5: the largest absolute Ez norm falls by more than 3.3 per halving of dx (FP64:
    at the finest grid the discretisation error, ~4e-3 of fields ~3e5, is below FP32 rounding)."""
    T = 40
    d = []
    for size, dx, steps in ((10, "0.001", T + 1), (20, "0.0005", 2 * T + 1), (40, "0.00025", 4 * T + 1)):
        args = (f"--3d --size x:{size} --same-size --angle-phi 0 --dx {dx} --wavelength 0.02 "
                "--use-sin1-border-condition --use-sin1-start-values --calc-sin1-diff-norm --eps-normed --mu-normed "
                f"--courant-factor 149896229 --time-steps {steps}")
        cfg = parse_args(args.split())
        cfg.precision = precision
        eng = Engine(cfg)
        eng.prepare()
        d.append(float(run_norms(eng, cfg)["Ez"][:, 0].max()))
        eng.close()
    say(f"[gpu] sin1 precision {precision}: {d}")
    assert d[0] / d[1] > 3.3 and d[1] / d[2] > 3.3


def test_cli_diff_norm_lines(engine_lib):
    """bin/fdtd3d with the t2.1 command line prints one "DIFF NORM Ez" / "DIFF NORM Hy" line per
    step in the reference's format; the suite's awk (the largest field 14) stays <= 1e-4."""
    exe = os.path.join(BIN, "fdtd3d")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    args = ("--time-steps 101 --size x:20 --same-size --3d --angle-phi 0 --dx 0.001 --wavelength 0.02 --log-level 0 "
            "--use-polinom1-border-condition --use-polinom1-start-values --use-polinom1-right-side "
            "--calc-polinom1-diff-norm --layout-type 0").split()
    say("[gpu] cli diff norm lines")
    out = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    ez = [ln for ln in out.stdout.splitlines() if "DIFF NORM Ez" in ln]
    hy = [ln for ln in out.stdout.splitlines() if "DIFF NORM Hy" in ln]
    assert len(ez) == 101 and len(hy) == 101
    assert ez[0].startswith("-> DIFF NORM Ez. Timestep 0. Value = ( ")
    assert ez[0].endswith(" %% )".replace("%%", "%"))
    worst = max(float(ln.split()[13]) for ln in ez)
    assert worst <= 1e-4
    # the same numbers as the engine's own
    cfg = parse_args(args)
    eng = Engine(cfg)
    eng.prepare()
    norms = run_norms(eng, cfg)
    eng.close()
    vals = np.array([[float(ln.split()[9]), float(ln.split()[13])] for ln in ez])
    assert np.allclose(vals, norms["Ez"][:, :2], rtol=1e-12, atol=1e-20)   # %.20f keeps 20 decimals


def _exp_engine(cfg):
    eng = Engine(cfg)
    eng.set_material_sphere("eps", cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius,
                            accuracy=cfg.sphere_accuracy, stair=cfg.stair_sphere)
    eng.prepare()
    return eng


def _exp_oracles(cfg, kinds):
    from oracle.binding import sphere_map
    parts = []
    for kind in kinds:
        o = Oracle(cfg)
        if kind:
            o.set_source_kind(kind)
        o.set_material("eps", sphere_map(cfg, cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius))
        o.prepare()
        parts.append(o)
    return parts


@pytest.mark.parametrize("args", EXP_CASES)
def test_exp_solutions_vs_oracle(args, engine_lib):
    """--calc-exp<n>-<pair>-diff-norm on the engine against the oracle (pinned bit for bit to the
    reference's real and complex builds): FP64, real and complex handles; fields within 1e-12
    (PML), every step's norms and percentages within 1e-8 relative (the device's sin / cos and
    its order of the sums; the norm is a difference ~1e-3 of the fields)."""
    import dataclasses
    from oracle.binding import oracle_exp_mod_norm
    from tests.helpers import run_norms_complex
    cfg = parse_args(args.split())
    say(f"[gpu] exp solution: {args[:12]} {args.split('--calc-')[1][:10]}")
    # real build
    (o,) = _exp_oracles(cfg, [0])
    ref = run_norms(o, cfg)
    eng = _exp_engine(cfg)
    got = run_norms(eng, cfg)
    assert sorted(got) == sorted(ref) and len(ref) == 2
    for n in field_names(cfg):
        assert rel_l2(eng.field(n), o.field(n)) <= 1e-12, n
    for n in ref:
        assert np.allclose(got[n][:, :2], ref[n][:, :2], rtol=1e-8, atol=0), n
    eng.close()
    o.close()
    # complex build
    ccfg = dataclasses.replace(cfg, complex_fields=True)
    re_, im_ = _exp_oracles(cfg, [1, 2])
    cref = run_norms_complex(re_, im_, cfg, oracle_exp_mod_norm)
    eng = _exp_engine(ccfg)
    cgot = {}
    for _ in range(cfg.time_steps):
        eng.step(1)
        for n in cref:
            v = eng.diff_norm_complex(n)
            cgot.setdefault(n, []).append((v[0], v[3], v[1], v[4], v[2], v[5]))
    for n in field_names(cfg):
        assert rel_l2(eng.field(n), re_.field(n)) <= 1e-12, n
        assert rel_l2(eng.imag().field(n), im_.field(n)) <= 1e-12, n
    for n in cref:
        assert np.allclose(np.array(cgot[n]), cref[n], rtol=1e-8, atol=0), n
    eng.close()


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("pair", ["exhy", "exhz", "eyhx", "eyhz", "ezhx", "ezhy"])
def test_suite_t7_plane_wave_and_dielectric(pair, layout, engine_lib):
    """Tests/suite t7.1 - t7.6 (one 1D pair each, complex build, 600 cells, dz 0.1 mm, lambda
    2 cm, PML 50, TF/SF border at 150 on the left only, the dielectric from 300, 10000 steps):
    exp1 in vacuum over [155, 545), exp3 inside eps = 4 over [305, 545), exp2 reflected over
    [55, 145); at the last step the real, imaginary and module percentages of both components
    stay under 0.12 / 0.12 / 0.17 %, in both Yee layouts (t7.x run --layout-type 0 and 1)."""
    import dataclasses
    from tests.helpers import exp_case
    for kind, norm, eps, bar in (("exp1", (155, 545), 1, 0.12), ("exp3", (305, 545), 4, 0.12),
                                 ("exp2", (55, 145), 4, 0.17)):
        args = exp_case(pair, kind, size=600, pml=50, tfsf=150, center=450, radius=150, steps=10000, dx="0.0001",
                        wavelength="0.02", eps=eps, norm=norm)
        cfg = dataclasses.replace(parse_args(args.split()), complex_fields=True, layout=layout)
        eng = _exp_engine(cfg)
        eng.step(cfg.time_steps)
        worst = 0.0
        for n in cfg.components:
            v = eng.diff_norm_complex(n)
            worst = max(worst, v[3], v[4], v[5])
        eng.close()
        say(f"[gpu] t7 {pair} {kind} layout {layout}: largest percentage {worst:.4f} (bar {bar})")
        assert 0 < worst <= bar, (pair, kind, worst)


def test_cli_complex_diff_norm_lines(engine_lib):
    """fdtd3d-complex prints the complex build's line, "-> DIFF NORM Ex. Timestep t. Value = ( re ,
    im ) = ( re % , im % ), module = m = ( m % )", whose fields 16, 19 and 27 the suite's awk
    reads; the numbers are the engine's.  The real binary prints the real line for the same run."""
    import dataclasses
    from tests.helpers import exp_case
    exe = os.path.join(BIN, "fdtd3d-complex")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    args = exp_case("exhy", "exp3", steps=120).split() + ["--log-level", "0"]
    say("[gpu] cli complex diff norm lines")
    out = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    ex = [ln for ln in out.stdout.splitlines() if "DIFF NORM Ex" in ln]
    hy = [ln for ln in out.stdout.splitlines() if "DIFF NORM Hy" in ln]
    assert len(ex) == 120 and len(hy) == 120
    last = ex[-1].split()
    assert ex[-1].startswith("-> DIFF NORM Ex. Timestep 119. Value = ( ") and last[20:22] == ["),", "module"]
    cfg = dataclasses.replace(parse_args(args), complex_fields=True)
    eng = _exp_engine(cfg)
    eng.step(cfg.time_steps)
    v = eng.diff_norm_complex("Ex")
    eng.close()
    printed = [float(last[i]) for i in (9, 11, 15, 18, 23, 26)]
    assert np.allclose(printed, [v[0], v[1], v[3], v[4], v[2], v[5]], rtol=1e-12, atol=1e-20)
    real = subprocess.run([os.path.join(BIN, "fdtd3d")] + args, capture_output=True, text=True, timeout=120)
    assert real.returncode == 0, real.stderr
    rl = [ln for ln in real.stdout.splitlines() if "DIFF NORM Ex" in ln]
    assert len(rl) == 120 and rl[-1].startswith("-> DIFF NORM Ex. Timestep 119. Value = ( ") and rl[-1].endswith(" % )")
