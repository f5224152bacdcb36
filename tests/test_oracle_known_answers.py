"""The C oracle against known answers the reference's own tests and the mathematics fix.

- unit-test-internalscheme.cpp:941-988: a 3D TF/SF plane wave in vacuum has |Ez| = 1
  inside the total-field box (to sqrt(1.8e-5)) and exactly 0 outside after 200 steps.
  The reference checks the magnitude of a complex run whose sources are sin + i cos;
  here it is assembled from a sin-source run and a cos-source run.
- Tests/suite/t6.1 and t6.7: the scattered field of a normally incident plane wave in vacuum
  (Examples/vacuum3D_scattered.txt, vacuum2D_TMz_scattered.txt) has magnitude <= 1e-14 with
  minimum 0; here it is exactly 0 for every component, which fixes the E/H incident grid,
  the half-cell shift of HInc, the projection and the inclusive box borders.
- unit-test-approximation.cpp:33-60: closed-form and Newton wave numbers agree to 1%.
- the closed-form wave numbers solve the Yee dispersion relation.
- the y/z PML sigma profile telescopes: sum of sigma over one slab = sigma_max * L / 7.
- mirror and transposition symmetry of a centred point source in TMz.
"""
import math

import numpy as np
import pytest

from fdtd3d_amd.config import Config, SCHEME_DIM3, SCHEME_TMZ, parse_args
from oracle.binding import Oracle, lib


def _run(cfg, phase=0.0, steps=None):
    o = Oracle(cfg)
    o.set_source_phase(phase)
    o.prepare()
    o.step(cfg.time_steps if steps is None else steps)
    return o


def test_unit_test_internalscheme_3d_tfsf_amplitude():
    # SIZE 20, TFSF 7, LAMBDA 0.2, DX 0.002, angles 90/90/90, --time-steps 200 --use-tfsf
    cfg = Config(scheme=SCHEME_DIM3, size=[20, 20, 20], pml=[5, 5, 5], tfsf_left=[7, 7, 7],
                 tfsf_right=[7, 7, 7], dx=0.002, wavelength=0.2, angle_deg=[90.0, 90.0, 90.0],
                 use_tfsf=True, time_steps=200)
    re = _run(cfg).field("Ez")
    im = _run(cfg, phase=math.pi / 2).field("Ez")
    mag = np.sqrt(re ** 2 + im ** 2)
    for i in range(20):
        for j in range(20):
            for k in range(20):
                x, y, z = i + 0.5, j + 0.5, k + 1.0   # minEzCoordFP
                if 7 <= x <= 13 and 7 <= y <= 13 and 7 <= z <= 13:
                    assert (mag[i, j, k] - 1.0) ** 2 < 0.000018, (i, j, k, mag[i, j, k])
                else:
                    assert mag[i, j, k] < 1e-7, (i, j, k, mag[i, j, k])


# Examples/vacuum3D_scattered.txt and vacuum2D_TMz_scattered.txt without the image options
VACUUM_SCATTERED = [
    ("--time-steps 100 --size x:40,y:40,z:40 --3d --angle-teta 0 --angle-phi 0 --angle-psi 0 --dx 0.0005 "
     "--wavelength 0.02 --use-tfsf --tfsf-size-left x:12,y:12,z:12 --tfsf-size-right x:12,y:12,z:12", "Ey"),
    ("--time-steps 100 --size x:40,y:40 --2d-tmz --angle-teta 90 --angle-phi 0 --angle-psi 90 --dx 0.0005 "
     "--wavelength 0.02 --use-tfsf --tfsf-size-left x:12,y:12 --tfsf-size-right x:12,y:12", "Ez"),
]


@pytest.mark.parametrize("args,main", VACUUM_SCATTERED)
def test_suite_t6_scattered_field_vanishes(args, main):
    cfg = parse_args(args.split())
    runs = [_run(cfg), _run(cfg, phase=math.pi / 2)]
    comps = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"] if cfg.scheme == SCHEME_DIM3 else ["Ez", "Hx", "Hy"]
    total = np.hypot(runs[0].field(main), runs[1].field(main))
    assert total.max() > 0.9   # the wave fills the total-field box
    for c in comps:
        scat = np.hypot(runs[0].scattered(c), runs[1].scattered(c))
        assert scat.max() <= 1e-14 and scat.min() == 0.0, (c, scat.max())


def test_wave_number_closed_form_vs_newton():
    L = lib()
    lam = 0.04
    step = math.pi / 12
    a1 = 0.0
    while a1 < 2 * math.pi + step:
        a2 = 0.0
        while a2 < 2 * math.pi + step:
            for courant in (0.5, 0.75, 1.0):
                delta = 0.0002
                while delta <= 0.002:
                    nl = lam / delta
                    g = L.oracle_wave_number_general(delta, lam, courant, nl, a1, a2)
                    k = L.oracle_wave_number(delta, lam, courant, nl, a1, a2)
                    assert (k - g) ** 2 < 2 * math.pi / lam / 100.0
                    delta *= 2
            a2 += step
        a1 += step


@pytest.mark.parametrize("a1,a2", [(0.0, 0.0), (math.pi / 2, 0.0), (math.pi / 2, math.pi / 4), (0.3, 1.1)])
def test_wave_number_solves_dispersion_relation(a1, a2):
    L = lib()
    dx, lam, S = 0.0005, 0.02, 0.5
    k = L.oracle_wave_number(dx, lam, S, lam / dx, a1, a2)
    kx, ky, kz = k * math.sin(a1) * math.cos(a2), k * math.sin(a1) * math.sin(a2), k * math.cos(a1)
    lhs = sum(math.sin(c * dx / 2) ** 2 for c in (kx, ky, kz))
    rhs = math.sin(math.pi * S / (lam / dx)) ** 2 / S ** 2
    assert abs(lhs - rhs) < 1e-9


def test_sigma_profiles_closed_form():
    P, N, dx = 6, 30, 0.0005
    cfg = Config(scheme=SCHEME_DIM3, size=[N, N, N], pml=[P, P, P], dx=dx, use_pml=True, time_steps=1)
    o = Oracle(cfg)
    o.prepare()
    eps0 = 1 / (4 * math.pi * 1e-7 * 299792458.0 ** 2)
    mu0 = 4 * math.pi * 1e-7
    L_ = P * dx
    smax = -math.log(1e-16) * 7 / (2 * math.sqrt(mu0 / eps0) * L_)
    sx = o.material(6)[:, 0, 0]
    sy = o.material(7)[0, :, 0]
    sz = o.material(8)[0, 0, :]
    # y/z: integer distances 0..P-1 telescope to sigma_max * L / 7
    assert math.isclose(sy[:P].sum() * dx, smax * L_ / 7, rel_tol=1e-12)
    assert math.isclose(sy[N - P:].sum() * dx, smax * L_ / 7, rel_tol=1e-12)
    np.testing.assert_array_equal(sy, sz)
    # x: fractional distances P-0.5 .. 0.5 on the left, 0.5 .. P-0.5 on the right
    f = smax / (dx * L_ ** 6 * 7)
    for i in range(P):
        d = P - (i + 0.5)
        assert math.isclose(sx[i], f * (((d + 1) * dx) ** 7 - (d * dx) ** 7), rel_tol=1e-12)
    assert np.all(sx[P:N - P] == 0) and np.all(sy[P:N - P] == 0)
    assert sx[N - P] > 0 and sy[N - P] > 0 and sy[N - P] < sx[N - P]


def test_sigma_follows_the_scheme_coordinate_slot():
    """Scheme::initSigmas (Scheme.cpp:141-420) applies initSigmaX to a scheme's first grid
    coordinate whatever physical axis it is: on the (y, z) plane of TMx / TEx the y profile is
    the fractional-distance one that x has in 3D, z keeps the truncated one."""
    P, N = 5, 24
    d3 = Oracle(Config(scheme=SCHEME_DIM3, size=[N, N, N], pml=[P, P, P], use_pml=True, time_steps=1))
    d3.prepare()
    sx3, sy3 = d3.material(6)[:, 0, 0], d3.material(7)[0, :, 0]
    for name in ("--2d-tmx", "--2d-tex"):
        cfg = parse_args(f"{name} --size y:{N},z:{N} --use-pml --pml-size y:{P},z:{P}".split())
        o = Oracle(cfg)
        o.prepare()
        sy = o.material(7).reshape(N, N)[:, 0]
        sz = o.material(8).reshape(N, N)[0, :]
        np.testing.assert_array_equal(sy, sx3)
        np.testing.assert_array_equal(sz, sy3)
        assert not np.array_equal(sy, sz)


def test_tmz_point_source_symmetry():
    n = 25
    c = n // 2
    cfg = parse_args(f"--2d-tmz --size x:{n},y:{n} --point-source x:{c},y:{c} --point-source-ez --time-steps 10".split())
    # the Yee stencil spreads one cell per step: after 10 steps no value has reached the
    # outer boundary, whose treatment is not mirror-symmetric (H at x = 1 is never updated)
    o = _run(cfg)
    ez, hx, hy = o.field("Ez"), o.field("Hx"), o.field("Hy")
    assert np.abs(ez).max() > 0
    np.testing.assert_allclose(ez, ez[::-1, :], rtol=0, atol=1e-12 * np.abs(ez).max())
    np.testing.assert_allclose(ez, ez[:, ::-1], rtol=0, atol=1e-12 * np.abs(ez).max())
    np.testing.assert_allclose(ez, ez.T, rtol=0, atol=1e-12 * np.abs(ez).max())
    # x <-> y exchange maps Hx onto -Hy (curl orientation)
    np.testing.assert_allclose(hx, -hy.T, rtol=0, atol=1e-12 * np.abs(hx).max())


@pytest.mark.parametrize("sigma", [0.0, 30.0])
@pytest.mark.parametrize("steps", [1, 37, 120])
def test_point_source_closed_form(sigma, steps):
    """A hard source assigns its value after the update of its component, so after the last
    step the source cell holds exactly sin(2 pi f dt ts) * exp(-((ts - t0)/sigma)^2) with
    ts = T - 1/2 (E) or T (H); sigma = 0 is the reference's plain sinusoid."""
    t0 = 60.0
    cfg = parse_args(f"--2d-tmz --size x:30,y:28 --use-pml --pml-size x:4,y:4 --point-source x:11,y:17 "
                     f"--point-source-ez --point-source-hy --time-steps {steps}".split())
    o = Oracle(cfg)
    o.set_source_gaussian(t0, sigma)
    o.prepare()
    o.step(steps)
    dt, f, _ = o.scalars()
    for name, ts in (("Ez", steps - 0.5), ("Hy", float(steps))):
        want = math.sin(dt * ts * 2 * math.pi * f)
        if sigma > 0:
            want *= math.exp(-((ts - t0) / sigma) ** 2)
        assert o.field(name)[11, 17] == want, name


def _unit_test_1d_vectors():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "unit_test_1d_vectors.json")
    with open(path) as fh:
        return json.load(fh)


def unit_test_1d_config(mode, v):
    """The 1D grid of the reference unit test (unit-test-internalscheme.cpp:33-38, test1D_<mode>):
    20 cells along the mode's axis, TF/SF 7 cells from each end, dx 0.002, lambda 0.2 (exact
    doubles, not CLI-rounded), 200 steps."""
    ax = v["axis"]
    t, p, s = v["angles"]
    cfg = parse_args(f"--1d-{mode.lower()} --size {ax}:20 --use-tfsf --tfsf-size-left {ax}:7 "
                     f"--tfsf-size-right {ax}:7 --angle-teta {t:g} --angle-phi {p:g} --angle-psi {s:g} "
                     "--time-steps 200".split())
    cfg.dx = 0.002
    cfg.wavelength = 0.2
    return cfg


@pytest.mark.parametrize("mode", ["ExHy", "ExHz", "EyHx", "EyHz", "EzHx", "EzHy"])
def test_unit_test_1d_golden_vectors(mode):
    """The reference's hand-written 1D golden vectors (unit-test-internalscheme.cpp:1033-1600,
    fixture written by tests/golden/extract_unit_test_vectors.py).  They come from a complex-field
    build whose plane-wave source is sin(arg) + i cos(arg) (InternalScheme.inc.h:3309-3313); every
    update coefficient is real, so the real part is the real-field run with the sin source: it must
    match bit for bit.  The imaginary part is the run with the cos source (the complex build's
    sources, oracle source kind 2): bit for bit as well.  The run with the sin source shifted by
    pi/2 (sin(arg + pi/2) rounds differently from cos(arg)) stays within the unit test's own bar,
    1e-7 (FPEXACT_ACCURACY, Kernels/FieldValue.h:150-160)."""
    v = _unit_test_1d_vectors()[mode]
    cfg = unit_test_1d_config(mode, v)
    re = Oracle(cfg)
    re.prepare()
    re.step(200)
    np.testing.assert_array_equal(re.field(v["component"]).ravel(), np.array(v["re"]))
    re1 = Oracle(cfg)
    re1.set_source_kind(1)
    re1.prepare()
    re1.step(200)
    np.testing.assert_array_equal(re1.field(v["component"]).ravel(), np.array(v["re"]))
    im = Oracle(cfg)
    im.set_source_kind(2)
    im.prepare()
    im.step(200)
    np.testing.assert_array_equal(im.field(v["component"]).ravel(), np.array(v["im"]))
    shifted = Oracle(cfg)
    shifted.set_source_phase(math.pi / 2)
    shifted.prepare()
    shifted.step(200)
    np.testing.assert_allclose(shifted.field(v["component"]).ravel(), np.array(v["im"]), rtol=0, atol=1e-7)
    # the TF/SF region carries a unit-amplitude wave, zero outside (test1D, :866-877)
    amp = np.hypot(re.field(v["component"]).ravel(), im.field(v["component"]).ravel())
    inside = amp > 0
    assert inside.sum() == 6 and np.all(np.abs(amp[inside] - 1) < math.sqrt(0.000018))


def test_sigma_x_closed_form_telescoping():
    """initSigmaX (SchemeHelper.h:597-634): the eps-point coordinate i + 1/2 against a fractional
    threshold, distance P - (i + 1/2) on the left and i + 1/2 - (N - P) on the right.  The cell
    values are the reference's polynomial grading sigma_max (x/L)^6 integrated over one cell
    (initSigma, :574-592), so over a slab they telescope: sum sigma dx = sigma_max dx
    ((P + 1/2)^7 - (1/2)^7) / (7 P^6), with the x-cells [0, P) and [N - P, N) nonzero, mirror
    images of each other.  A whole-cell distance (P - i) would give (P + 1)^7 - 1, a threshold of
    N + 1/2 - P (the y / z rule) P^7: both fail here."""
    for P, N, dx in ((6, 30, 0.0005), (16, 64, 0.000625), (3, 11, 0.002)):
        cfg = Config(scheme=SCHEME_DIM3, size=[N, N, N], pml=[P, P, P], dx=dx, use_pml=True, time_steps=1)
        o = Oracle(cfg)
        o.prepare()
        eps0 = 1 / (4 * math.pi * 1e-7 * 299792458.0 ** 2)
        mu0 = 4 * math.pi * 1e-7
        smax = -math.log(1e-16) * 7 / (2 * math.sqrt(mu0 / eps0) * (P * dx))
        sx = o.material(6)[:, 0, 0]
        want = smax * dx * ((P + 0.5) ** 7 - 0.5 ** 7) / (7 * P ** 6)
        assert math.isclose(sx[:P].sum() * dx, want, rel_tol=1e-12)
        assert math.isclose(sx[N - P:].sum() * dx, want, rel_tol=1e-12)
        assert np.all(sx[:P] > 0) and np.all(sx[N - P:] > 0) and np.all(sx[P:N - P] == 0)
        np.testing.assert_allclose(sx, sx[::-1], rtol=1e-13, atol=0)
        assert np.all(np.diff(sx[:P]) < 0)
