"""HIP engine parity against the oracle and the reference fixtures (needs the GPU).

Bars (per component, after the last step):
  - FP64, runs without PML: bit-identical to the reference (same operations, no FMA);
  - FP64 with PML / Drude: relative L2 <= 1e-12 (the engine updates E/H directly outside
    the PML slabs, which equals the reference's D->E stage with sigma = 0 up to rounding);
  - FP32: relative L2 of the E and H vectors <= 1e-5 against the FP64 oracle, and of every
    component holding at least 10 % of its vector's squared norm against its own norm (the
    smaller components are scattered / cross-polarised fields, on which the FP32 rounding of
    the large ones weighs more);
  - TF/SF and PML masks: bit-exact.
"""
import sys

import numpy as np
import pytest

from fdtd3d_amd.config import parse_args
from fdtd3d_amd.engine import Engine, run_engine
from oracle.binding import Oracle, run_oracle, sphere_map
from tests.helpers import check_fp32, field_names, golden_names, load_golden, rel_l2

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_FP32 = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def vec_rel(eng_fields, ref_fields):
    num = sum(np.sum((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2) for a, b in zip(eng_fields, ref_fields))
    den = sum(np.sum(np.asarray(b, np.float64) ** 2) for b in ref_fields)
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))




@pytest.mark.parametrize("name", golden_names())
def test_engine_fp64_vs_reference_fixture(name, engine_lib):
    cfg, mats, data = load_golden(name)
    say(f"[gpu] fixture {name} fp64")
    eng = run_engine(cfg, mats)
    exact = not cfg.use_pml
    # a fixture of the complex-field build holds the imaginary parts as "<name>.im"
    parts = [(eng, "")] + ([(eng.imag(), ".im")] if cfg.complex_fields else [])
    for part, sfx in parts:
        for f in field_names(cfg):
            got = part.field(f).ravel()
            ref = data[f + sfx]
            if exact:
                assert np.array_equal(got, ref), f"{name}/{f}{sfx}: max diff {np.abs(got - ref).max()}"
            else:
                assert rel_l2(got, ref) <= TOL_FP64, f"{name}/{f}{sfx}: rel {rel_l2(got, ref)}"
        if cfg.use_tfsf:
            e, h = part.incident(256)
            np.testing.assert_array_equal(e, data["EInc" + sfx])
            np.testing.assert_array_equal(h, data["HInc" + sfx])
    eng.close()


@pytest.mark.parametrize("name", golden_names())
def test_engine_fp32_vs_fp64_oracle(name, engine_lib):
    cfg, mats, data = load_golden(name)
    cfg.precision = 4
    say(f"[gpu] fixture {name} fp32")
    eng = run_engine(cfg, mats)
    check_fp32(name, field_names(cfg), eng.field, lambda n: data[n].reshape(cfg.shape))
    if cfg.complex_fields:
        im = eng.imag()
        check_fp32(name + " (imaginary part)", field_names(cfg), im.field, lambda n: data[n + ".im"].reshape(cfg.shape))
    eng.close()


MASK_CASES = [
    "--3d --size x:17,y:15,z:19 --use-tfsf --tfsf-size-left x:4,y:3,z:5 --tfsf-size-right x:5,y:4,z:3 "
    "--angle-teta 30 --angle-phi 45 --angle-psi 90 --time-steps 1",
    "--2d-tmz --size x:23,y:29 --use-tfsf --tfsf-size-left x:5,y:6 --tfsf-size-right x:4,y:7 "
    "--angle-teta 90 --angle-phi 30 --angle-psi 90 --time-steps 1",
    "--2d-tez --size x:23,y:29 --use-tfsf --tfsf-size-left x:5,y:6 --tfsf-size-right x:4,y:7 "
    "--angle-teta 90 --angle-phi 30 --angle-psi 0 --time-steps 1",
    "--2d-tmx --size x:9,y:23,z:27 --use-tfsf --tfsf-size-left y:5,z:6 --tfsf-size-right y:4,z:7 "
    "--angle-teta 35 --angle-phi 90 --angle-psi 0 --time-steps 1",
    "--2d-tey --size x:26,y:9,z:21 --use-tfsf --tfsf-size-left x:5,z:6 --tfsf-size-right x:6,z:4 "
    "--angle-teta 50 --angle-phi 0 --angle-psi 90 --time-steps 1",
]


@pytest.mark.parametrize("args", MASK_CASES)
def test_tfsf_masks_bit_exact(args, engine_lib):
    cfg = parse_args(args.split())
    say(f"[gpu] masks {args[:40]}")
    o = Oracle(cfg)
    o.prepare()
    eng = Engine(cfg)
    eng.prepare()
    for comp in field_names(cfg):
        np.testing.assert_array_equal(eng.tfsf_mask(comp), o.tfsf_mask(comp), err_msg=comp)
        assert o.tfsf_mask(comp).any()
    eng.close()


LARGER = [
    # config 1 geometry at half size: sphere, PML, TF/SF, 120 steps
    ("cfg1_half", "--3d --size x:40 --same-size --dx 0.0005 --wavelength 0.02 --angle-teta 0 --angle-phi 0 "
     "--angle-psi 0 --use-tfsf --tfsf-size-left x:10 --tfsf-size-right x:10 --same-size-tfsf --use-pml "
     "--pml-size x:5 --same-size-pml --courant-factor 0.5 --time-steps 120 --eps-sphere 4 "
     "--eps-sphere-center x:20,y:20,z:20 --eps-sphere-radius 10"),
    # config 2 features at small size: oblique incidence, block eps in [1, 12]
    ("cfg2_small", "--3d --size x:48,y:40,z:56 --use-pml --pml-size x:6 --same-size-pml --use-tfsf "
     "--tfsf-size-left x:10 --tfsf-size-right x:10 --same-size-tfsf --angle-teta 30 --angle-phi 45 "
     "--angle-psi 90 --time-steps 100"),
    # TMz point source with PML and random eps blobs
    ("tmz_blobs", "--2d-tmz --size x:96,y:80 --use-pml --pml-size x:12,y:12 --point-source x:40,y:33 "
     "--point-source-ez --time-steps 200"),
    # rows of 200 cells: the fused kernel's box holds 187 z cells of each row, three waves with
    # the z+1 neighbours of each wave's last lane passed through LDS; oblique TF/SF, block eps
    ("fused_rows_3_waves", "--3d --size x:24,y:22,z:200 --use-pml --pml-size x:6 --same-size-pml --use-tfsf "
     "--tfsf-size-left x:8,y:8,z:40 --tfsf-size-right x:8,y:8,z:40 --angle-teta 30 --angle-phi 45 "
     "--angle-psi 90 --time-steps 120"),
]


def _materials(name, cfg):
    rng = np.random.default_rng(20251202)
    mats = {}
    if cfg.eps_sphere != 1.0:
        mats["eps"] = sphere_map(cfg, cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius)
    if name in ("cfg2_small", "fused_rows_3_waves"):
        nb = [(s + 7) // 8 for s in cfg.shape]
        blocks = rng.uniform(1.0, 12.0, size=nb)
        mats["eps"] = np.kron(blocks, np.ones((8, 8, 8)))[: cfg.shape[0], : cfg.shape[1], : cfg.shape[2]]
    if name == "tmz_blobs":
        x, y = np.meshgrid(np.arange(cfg.shape[0]), np.arange(cfg.shape[1]), indexing="ij")
        eps = np.ones(cfg.shape)
        for _ in range(5):
            cx, cy, r = rng.uniform(15, 80), rng.uniform(15, 65), rng.uniform(3, 8)
            eps += 8 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * r * r))
        mats["eps"] = np.clip(eps, 1, 9)
    return mats


@pytest.mark.parametrize("name,args", LARGER)
@pytest.mark.parametrize("precision", [8, 4])
def test_engine_vs_oracle_larger(name, args, precision, engine_lib):
    cfg = parse_args(args.split())
    mats = _materials(name, cfg)
    o = run_oracle(cfg, mats)
    cfg.precision = precision
    say(f"[gpu] larger {name} precision {precision}")
    eng = run_engine(cfg, mats)
    names = field_names(cfg)
    if precision == 8:
        for f in names:
            r = rel_l2(eng.field(f), o.field(f))
            assert r <= TOL_FP64, f"{name}/{f}: {r}"
    else:
        check_fp32(name, names, eng.field, o.field)
    eng.close()


SCATTERED_CASES = [
    "--3d --size x:30,y:26,z:28 --use-pml --pml-size x:5 --same-size-pml --use-tfsf --tfsf-size-left x:8,y:7,z:9 "
    "--tfsf-size-right x:9,y:8,z:7 --angle-teta 30 --angle-phi 40 --angle-psi 20 --time-steps 60",
    "--2d-tmz --size x:40,y:36 --use-pml --pml-size x:6,y:6 --use-tfsf --tfsf-size-left x:10,y:9 "
    "--tfsf-size-right x:9,y:10 --angle-teta 90 --angle-phi 30 --angle-psi 90 --time-steps 80",
    "--2d-tez --size x:40,y:36 --use-pml --pml-size x:6,y:6 --use-tfsf --tfsf-size-left x:10,y:9 "
    "--tfsf-size-right x:9,y:10 --angle-teta 90 --angle-phi 50 --angle-psi 0 --time-steps 80",
]


@pytest.mark.parametrize("args", SCATTERED_CASES)
@pytest.mark.parametrize("precision", [8, 4])
def test_scattered_field_vs_oracle(args, precision, engine_lib):
    """yee_get_scattered_field against the oracle's makeGridScattered restatement (oblique
    incidence, so the subtracted wave is interpolated and projected on every component)."""
    cfg = parse_args(args.split())
    o = run_oracle(cfg)
    cfg.precision = precision
    say(f"[gpu] scattered {cfg.scheme} precision {precision}")
    eng = run_engine(cfg)
    for f in field_names(cfg):
        ref = o.scattered(f)
        got = eng.scattered(f)
        if precision == 8:
            assert rel_l2(got, ref) <= TOL_FP64, f"{f}: {rel_l2(got, ref)}"
        else:
            # scattered = total - incident: each term carries the FP32 bar relative to its own
            # size, so the bar scales with (|total| + |incident|) / |scattered|
            tot = o.field(f)
            scale = (np.linalg.norm(tot) + np.linalg.norm(tot - ref)) / np.linalg.norm(ref)
            assert rel_l2(got, ref) <= TOL_FP32 * scale, f"{f}: {rel_l2(got, ref)} (scale {scale:.2f})"
    eng.close()


CURRENT_CASES = [
    # J / M current sources (Scheme.inc.h:1171-1260) with PML (subtracted from D) and TF/SF
    "--3d --size x:26,y:24,z:28 --use-pml --pml-size x:4 --same-size-pml --use-tfsf --tfsf-size-left x:7 "
    "--tfsf-size-right x:7 --same-size-tfsf --angle-teta 30 --angle-phi 45 --angle-psi 90 "
    "--current-source x:13,y:12,z:14 --current-source-jz --current-source-my --point-source x:13,y:12,z:14 "
    "--point-source-ex --time-steps 50",
    # no PML: the direct update, a current on a TF/SF face cell (FP32 bar 2e-5: see CURRENT_FP32_TOL)
    "--3d --size x:20,y:18,z:22 --use-tfsf --tfsf-size-left x:5 --tfsf-size-right x:5 --same-size-tfsf "
    "--angle-teta 0 --angle-phi 0 --angle-psi 0 --current-source x:4,y:9,z:11 --current-source-jx "
    "--current-source-mz --time-steps 40",
    # in a PML slab, Drude metamaterials on
    "--3d --size x:22,y:20,z:24 --use-pml --pml-size x:4 --same-size-pml --use-metamaterials "
    "--current-source x:2,y:10,z:12 --current-source-jy --current-source-mx --time-steps 40",
    "--2d-tmz --size x:36,y:30 --use-pml --pml-size x:5,y:5 --current-source x:17,y:14 --current-source-jz "
    "--current-source-mx --time-steps 60",
    "--2d-tey --size x:30,y:5,z:28 --use-pml --pml-size x:4,z:4 --current-source x:14,y:0,z:13 "
    "--current-source-jx --current-source-my --angle-phi 0 --time-steps 60",
]


# FP32 bars of the current-source cases.  Next to a J source without PML, E reaches ~7e4 V/m
# beside H ~ 1 A/m, and the curl of E that updates H cancels most of it: the reference's own
# FP32 build (refdrv_f32_fast) deviates from FP64 by 1.24e-5 in H on case 1
# (tests/test_oracle_ref_live.py::test_reference_fp32_deviation_next_to_current_source), so its bar
# is 2e-5; the others keep 1e-5.
CURRENT_FP32_TOL = {1: 2e-5}


@pytest.mark.parametrize("case", range(len(CURRENT_CASES)))
@pytest.mark.parametrize("precision", [8, 4])
def test_current_sources_vs_oracle(case, precision, engine_lib):
    args = CURRENT_CASES[case]
    cfg = parse_args(args.split())
    o = run_oracle(cfg)
    cfg.precision = precision
    say(f"[gpu] current sources precision {precision}: {args[:50]}")
    eng = run_engine(cfg)
    names = field_names(cfg)
    assert max(float(np.abs(o.field(f)).max()) for f in names) > 0
    if precision == 8:
        for f in names:
            got, want = eng.field(f), o.field(f)
            if cfg.use_pml:
                assert rel_l2(got, want) <= TOL_FP64, f"{f}: {rel_l2(got, want)}"
            else:
                assert np.array_equal(got, want), f"{f}: max diff {np.abs(got - want).max()}"
    else:
        for group in ("E", "H"):
            ns = [n for n in names if n[0] == group]
            r = vec_rel([eng.field(n) for n in ns], [o.field(n) for n in ns])
            assert r <= CURRENT_FP32_TOL.get(case, TOL_FP32), f"{group}: {r}"
    eng.close()


@pytest.mark.parametrize("args,box", [
    ("--2d-tmz --size x:40,y:36 --use-pml --pml-size x:6,y:6 --use-metamaterials --current-source x:20,y:18 "
     "--current-source-jz --current-source-mx --time-steps 150", (slice(15, 25), slice(13, 23))),
    ("--3d --size x:30,y:28,z:32 --use-pml --pml-size x:4,y:4,z:4 --use-metamaterials --current-source x:15,y:14,z:16 "
     "--current-source-jx --current-source-mz --time-steps 80", (slice(12, 19), slice(11, 18), slice(13, 20))),
])
@pytest.mark.parametrize("precision", [8, 4])
def test_current_source_inside_drude_box(args, box, precision, engine_lib):
    """A J / M source inside a resonant Drude region: the current enters D / B before the ADE
    stage (FP32: the increment form of the ADE carries it as -jdx*cbc in dD)."""
    cfg = parse_args(args.split())
    w0 = 2 * np.pi * 299792458.0 / cfg.wavelength
    om, ga = np.zeros(cfg.shape), np.zeros(cfg.shape)
    om[box], ga[box] = 1.5 * w0, 0.01 * w0
    mats = {"omegape": om, "omegapm": om, "gammae": ga, "gammam": ga}
    o = run_oracle(cfg, mats)
    cfg.precision = precision
    say(f"[gpu] current source in Drude box precision {precision}: {args[:40]}")
    eng = run_engine(cfg, mats)
    names = field_names(cfg)
    if precision == 8:
        for f in names:
            assert rel_l2(eng.field(f), o.field(f)) <= TOL_FP64, f
    else:
        check_fp32("current source in Drude box", names, eng.field, o.field)
    eng.close()


def test_special_mask_covers_every_nonzero_sigma(engine_lib):
    cfg = parse_args("--3d --size x:20,y:18,z:22 --use-pml --pml-size x:4,y:3,z:5 --time-steps 1".split())
    say("[gpu] special mask")
    eng = Engine(cfg)
    eng.prepare()
    special = eng.special_mask().astype(bool)
    o = Oracle(cfg)
    o.prepare()
    sig = [o.material(6 + a) for a in range(3)]
    # a field cell averages eps cells p and p+1 along each axis
    any_sigma = np.zeros(cfg.shape, bool)
    for s in sig:
        nz = s != 0
        grown = nz.copy()
        grown[:-1, :, :] |= nz[1:, :, :]
        grown[:, :-1, :] |= nz[:, 1:, :]
        grown[:, :, :-1] |= nz[:, :, 1:]
        any_sigma |= grown
    assert np.all(special[any_sigma])
    eng.close()


SPLIT_VS_FUSED = [
    # PML + TF/SF oblique + point source inside the PML-free box, eps blocks
    ("--3d --size x:40,y:36,z:44 --use-pml --pml-size x:5,y:4,z:6 --use-tfsf --tfsf-size-left x:9,y:8,z:10 "
     "--tfsf-size-right x:8,y:9,z:9 --angle-teta 30 --angle-phi 45 --angle-psi 90 --point-source x:20,y:17,z:22 "
     "--point-source-ez --point-source-hx --time-steps 60"),
    # no PML (the box is the whole interior), point source in a z boundary lane of the box
    ("--3d --size x:24,y:30,z:70 --use-tfsf --tfsf-size-left x:5,y:6,z:7 --tfsf-size-right x:6,y:5,z:8 "
     "--angle-teta 60 --angle-phi 10 --angle-psi 30 --point-source x:12,y:15,z:1 --point-source-ey "
     "--time-steps 40"),
    # point source inside a PML slab, nz not a multiple of 64
    ("--3d --size x:30,y:26,z:90 --use-pml --pml-size x:4,y:3,z:7 --point-source x:2,y:13,z:45 "
     "--point-source-hz --point-source-ex --time-steps 50"),
    # current sources J / M inside the fused box and inside a PML slab
    ("--3d --size x:40,y:36,z:44 --use-pml --pml-size x:5,y:4,z:6 --current-source x:20,y:17,z:22 "
     "--current-source-jz --current-source-mx --time-steps 40"),
    ("--3d --size x:30,y:26,z:40 --use-pml --pml-size x:4,y:3,z:6 --current-source x:15,y:2,z:20 "
     "--current-source-jx --current-source-my --time-steps 40"),
    # rows of 10 waves (the 12-wave instantiation of the fused kernel), PML + TF/SF
    ("--3d --size x:14,y:12,z:620 --use-pml --pml-size x:3,y:2,z:5 --use-tfsf --tfsf-size-left x:5,y:4,z:6 "
     "--tfsf-size-right x:5,y:4,z:7 --angle-teta 20 --angle-phi 70 --angle-psi 10 --time-steps 30"),
]


@pytest.mark.parametrize("args", SPLIT_VS_FUSED)
@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("block", [4, 1])
def test_fused_step_matches_half_steps(args, precision, block, engine_lib, monkeypatch):
    """The fused E+H kernel (PML-free box) and the half-step kernels compute the same
    expressions in the same order: results, D / B included, must be bit-identical.  eps is
    constant on 4^3 blocks or random per cell (block 1: every z+1 / y+1 / x+1 average of the
    kernel's eps exchange differs from its neighbour's)."""
    cfg = parse_args(args.split())
    cfg.precision = precision
    rng = np.random.default_rng(7)
    nb = [(s + block - 1) // block for s in cfg.shape]
    eps = np.kron(rng.uniform(1.0, 6.0, size=nb), np.ones((block,) * 3))[: cfg.shape[0], : cfg.shape[1], : cfg.shape[2]]
    mats = {"eps": eps}
    say(f"[gpu] split vs fused precision {precision} eps block {block}: {args[:60]}")
    monkeypatch.setenv("YEE_KERNEL", "split")
    a = run_engine(cfg, mats)
    monkeypatch.setenv("YEE_KERNEL", "fused")
    b = run_engine(cfg, mats)
    for f in field_names(cfg, aux=True):
        fa, fb = a.field_raw(f), b.field_raw(f)
        assert np.array_equal(fa, fb), f"{f}: max diff {np.max(np.abs(fa.astype(np.float64) - fb))}"
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [8, 4])
def test_fused_step_with_drude_box_matches_half_steps(precision, engine_lib, monkeypatch):
    """Drude runs: the fused kernel leaves the Drude box to the half-step kernels; bit-identical
    to the all-half-step run."""
    cfg = parse_args("--3d --size x:36,y:32,z:40 --use-pml --pml-size x:4,y:4,z:5 --use-tfsf "
                     "--tfsf-size-left x:8,y:8,z:9 --tfsf-size-right x:8,y:8,z:9 --angle-teta 0 --angle-phi 0 "
                     "--angle-psi 0 --use-metamaterials --time-steps 50".split())
    cfg.precision = precision
    w0 = 2 * np.pi * 299792458.0 / cfg.wavelength
    om = np.zeros(cfg.shape)
    ga = np.zeros(cfg.shape)
    om[14:22, 12:20, 16:25] = 1.5 * w0
    ga[14:22, 12:20, 16:25] = 0.01 * w0
    mats = {"omegape": om, "omegapm": om, "gammae": ga, "gammam": ga}
    say(f"[gpu] drude split vs fused precision {precision}")
    monkeypatch.setenv("YEE_KERNEL", "split")
    a = run_engine(cfg, mats)
    monkeypatch.setenv("YEE_KERNEL", "fused")
    b = run_engine(cfg, mats)
    for f in field_names(cfg, aux=True):
        fa, fb = a.field_raw(f), b.field_raw(f)
        assert np.array_equal(fa, fb), f"{f}: max diff {np.max(np.abs(fa.astype(np.float64) - fb))}"
    a.close()
    b.close()


def test_fp32_drude_resonance_long_run(engine_lib):
    """A resonant Drude cube (omega_p = 1.5 omega, gamma = 0.01 omega) driven for 500 steps: the
    FP32 run stays within the FP32 bar of the FP64 oracle (the FP32 Drude cells carry the ADE in
    increments; in the reference's form the second difference of FP32 D drifts by ~2e-5 a step)."""
    cfg = parse_args("--3d --size x:24,y:24,z:24 --use-pml --pml-size x:5,y:5,z:5 --use-tfsf --tfsf-size-left x:7 "
                     "--tfsf-size-right x:7 --same-size-tfsf --angle-teta 0 --angle-phi 0 --angle-psi 0 "
                     "--use-metamaterials --time-steps 500".split())
    w0 = 2 * np.pi * 299792458.0 / cfg.wavelength
    om = np.zeros(cfg.shape)
    ga = np.zeros(cfg.shape)
    om[9:15, 9:15, 9:15] = 1.5 * w0
    ga[9:15, 9:15, 9:15] = 0.01 * w0
    mats = {"omegape": om, "omegapm": om, "gammae": ga, "gammam": ga}
    o = run_oracle(cfg, mats)
    cfg.precision = 4
    say("[gpu] fp32 drude resonance 500 steps")
    eng = run_engine(cfg, mats)
    check_fp32("drude resonance 500 steps", field_names(cfg), eng.field, o.field)
    eng.close()


COMPLEX_CASES = [
    # oblique TF/SF, eps blocks, point and current sources, PML; three-wave fused rows
    "--3d --size x:24,y:22,z:150 --use-pml --pml-size x:5 --same-size-pml --use-tfsf --tfsf-size-left x:7,y:7,z:30 "
    "--tfsf-size-right x:7,y:7,z:30 --angle-teta 30 --angle-phi 45 --angle-psi 90 --point-source x:12,y:11,z:70 "
    "--point-source-hy --current-source x:10,y:12,z:80 --current-source-jx --current-source-mz --time-steps 80",
    # Drude cube inside PML + TF/SF
    "--3d --size x:24,y:24,z:24 --use-pml --pml-size x:4 --same-size-pml --use-tfsf --tfsf-size-left x:7 "
    "--tfsf-size-right x:7 --same-size-tfsf --angle-teta 0 --angle-phi 0 --angle-psi 0 --use-metamaterials "
    "--time-steps 80",
    # TMz with TF/SF and sources
    "--2d-tmz --size x:60,y:52 --use-pml --pml-size x:8,y:8 --use-tfsf --tfsf-size-left x:14,y:13 "
    "--tfsf-size-right x:13,y:14 --angle-teta 90 --angle-phi 30 --angle-psi 90 --current-source x:30,y:26 "
    "--current-source-jz --point-source x:20,y:18 --point-source-ez --time-steps 100",
]


@pytest.mark.parametrize("args", COMPLEX_CASES)
@pytest.mark.parametrize("precision", [8, 4])
def test_complex_fields_vs_oracle(args, precision, engine_lib):
    """Complex field values (YEE_COMPLEX_FIELDS): both parts against the oracle's runs with the
    complex build's sources (pinned bit for bit to the reference's complex build)."""
    import dataclasses
    from oracle.binding import run_oracle_complex
    cfg = dataclasses.replace(parse_args(args.split()), complex_fields=True)
    mats = {}
    if "--use-metamaterials" in args:
        w0 = 2 * np.pi * 299792458.0 / cfg.wavelength
        om, ga = np.zeros(cfg.shape), np.zeros(cfg.shape)
        om[9:15, 9:15, 9:15], ga[9:15, 9:15, 9:15] = 1.5 * w0, 0.01 * w0
        mats = {"omegape": om, "omegapm": om, "gammae": ga, "gammam": ga}
    elif cfg.dim == 3:
        rng = np.random.default_rng(5)
        nb = [(s + 7) // 8 for s in cfg.shape]
        mats["eps"] = np.kron(rng.uniform(1.0, 6.0, size=nb), np.ones((8, 8, 8)))[: cfg.shape[0], : cfg.shape[1], : cfg.shape[2]]
    re_, im_ = run_oracle_complex(cfg, mats)
    cfg.precision = precision
    say(f"[gpu] complex fields precision {precision}: {args[:50]}")
    eng = run_engine(cfg, mats)
    im = eng.imag()
    for label, part, ref in (("real", eng, re_), ("imag", im, im_)):
        if precision == 8:
            for f in field_names(cfg):
                r = rel_l2(part.field(f), ref.field(f))
                assert r <= TOL_FP64, f"{label} {f}: {r}"
        else:
            check_fp32(f"complex {label}", field_names(cfg), part.field, ref.field)
    eng.close()


@pytest.mark.parametrize("args", [MASK_CASES[0], MASK_CASES[1]])
def test_h_centered_masks_are_the_role_swapped_e_centered_masks(args, engine_lib):
    """H_CENTERED puts each E component where E_CENTERED puts the H component of its axis, and its
    TF/SF predicates are the ones of that position (YeeGridLayout.cpp:292-306 for Ex against
    :1130-1147 for Hx): the H_CENTERED mask of Ex equals the E_CENTERED mask of Hx bit for bit, and
    so on for every pair of the scheme (3D, and TMz on the 3D kernels)."""
    import dataclasses
    cfg0 = parse_args(args.split())
    cfg1 = dataclasses.replace(cfg0, layout=1)
    say(f"[gpu] H_CENTERED masks {args[:40]}")
    e0, e1 = Engine(cfg0), Engine(cfg1)
    e0.prepare()
    e1.prepare()
    swap = {"Ex": "Hx", "Ey": "Hy", "Ez": "Hz", "Hx": "Ex", "Hy": "Ey", "Hz": "Ez"}
    for comp in field_names(cfg0):
        if swap[comp] not in field_names(cfg0):
            continue
        m1 = e1.tfsf_mask(comp)
        np.testing.assert_array_equal(m1, e0.tfsf_mask(swap[comp]), err_msg=comp)
        assert m1.any()
    e0.close()
    e1.close()


@pytest.mark.parametrize("layout", [0, 1])
def test_scattered_field_vanishes_in_vacuum(layout, engine_lib):
    """Without a scatterer the total field inside the TF/SF box is the incident wave: total minus
    the interpolated incident value at each component's position stays at the discretisation
    residue (a position off by half a cell leaves ~k dx / 2 of the amplitude)."""
    import dataclasses
    args = ("--3d --size x:40,y:36,z:44 --use-pml --pml-size x:6 --same-size-pml --use-tfsf "
            "--tfsf-size-left x:10 --tfsf-size-right x:10 --same-size-tfsf --angle-teta 30 --angle-phi 45 "
            "--angle-psi 90 --dx 0.0005 --wavelength 0.02 --time-steps 120")
    cfg = dataclasses.replace(parse_args(args.split()), layout=layout)
    say(f"[gpu] scattered field in vacuum, layout {layout}")
    eng = Engine(cfg)
    eng.prepare()
    eng.step(cfg.time_steps)
    eng.sync()
    for comp in ("Ex", "Ey", "Ez"):
        tot = eng.field(comp)
        sc = eng.scattered(comp)
        inner = (slice(12, 28), slice(12, 24), slice(12, 32))
        ratio = np.abs(sc[inner]).max() / np.abs(tot[inner]).max()
        say(f"[gpu]   {comp}: max |scattered| / max |total| inside the box = {ratio:.2e}")
        assert np.abs(tot[inner]).max() > 0.1
        assert ratio < 0.02, (comp, ratio)
    eng.close()
