"""Shared helpers of the test suite: fixture loading and parity metrics."""
from __future__ import annotations

import glob
import os
import shlex

import numpy as np

from fdtd3d_amd.config import parse_args

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

FIELDS_3D = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]
FIELDS_TMZ = ["Ez", "Hx", "Hy"]
FIELDS_TEZ = ["Ex", "Ey", "Hz"]


# fixtures that hold a summary of a run too large to store whole (see tests/golden/make_golden.py)
SUMMARY_FIXTURES = ("cfg1_sphere80",)


def golden_names():
    """Fixtures that hold every field of their run in full."""
    return sorted(n for n in (os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
                  if n not in SUMMARY_FIXTURES)


def load_summary(name):
    """(config, fixture dict) of a summary fixture: per component the L2 norm, sum, largest
    magnitude, mid-z plane and a seeded sample of cells ("cells": flat indices), and the eps grid."""
    data = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    return parse_args(shlex.split(str(data["argv"]))), data


def summary_errors(fields: dict, data: dict, shape) -> dict:
    """Relative errors of fields (name -> array of the grid's shape) against a summary fixture:
    L2 norm, sum (relative to the L2 norm times sqrt(cells)), largest magnitude, mid-z plane
    (relative L2) and the sampled cells (relative to the largest magnitude)."""
    out = {}
    n = float(np.prod(shape))
    for c, v in fields.items():
        v = np.asarray(v, dtype=np.float64).reshape(shape)
        l2 = float(data[c + "_l2"])
        e = {"l2": abs(float(np.sqrt(np.sum(v * v))) - l2) / l2,
             "sum": abs(float(np.sum(v)) - float(data[c + "_sum"])) / (l2 * np.sqrt(n)),
             "max": abs(float(np.abs(v).max()) - float(data[c + "_max"])) / float(data[c + "_max"]),
             "midz": rel_l2(v[:, :, shape[2] // 2], data[c + "_midz"]),
             "cells": float(np.abs(v.ravel()[data["cells"]] - data[c + "_cells"]).max()) / float(data[c + "_max"])}
        out[c] = e
    return out


def load_golden(name):
    """(config, materials, fixture dict) of one golden fixture."""
    data = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    argv = shlex.split(str(data["argv"]))
    # material files named in argv are reproduced from the arrays stored in the fixture
    clean = []
    skip = False
    for a in argv:
        if skip:
            skip = False
            continue
        if a.startswith("--load-") and a.endswith("-from-file"):
            skip = True
            continue
        clean.append(a)
    cfg = parse_args(clean)
    cfg.complex_fields = bool(data.get("complex", False))   # written by the complex-field build
    shape = cfg.shape
    mats = {}
    if not np.all(data["Eps"] == 1.0):
        mats["eps"] = data["Eps"].reshape(shape)
    for key, name_ in (("OmegaPE", "omegape"), ("OmegaPM", "omegapm"), ("GammaE", "gammae"), ("GammaM", "gammam")):
        if key in data:
            mats[name_] = data[key].reshape(shape)
    return cfg, mats, data


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    if nb == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)


def field_names(cfg, aux=False):
    base = list(cfg.components)
    names = list(base)
    if aux and cfg.use_pml:
        names += [("D" if n[0] == "E" else "B") + n[1] for n in base]
    if aux and cfg.use_metamaterials:
        names += [("D1" if n[0] == "E" else "B1") + n[1] for n in base]
    return names


def check_fp32(label, names, eng_field, ref_field, tol=1e-5):
    """FP32 bar: the E and H vectors, and each component with >= 10 % of its vector's squared
    norm, within tol relative L2 of the FP64 reference; every component's error printed."""
    for group in ("E", "H"):
        ns = [n for n in names if n[0] == group]
        if not ns:
            continue
        got = {n: np.asarray(eng_field(n), np.float64) for n in ns}
        ref = {n: np.asarray(ref_field(n), np.float64) for n in ns}
        num = {n: float(np.sum((got[n] - ref[n]) ** 2)) for n in ns}
        den = {n: float(np.sum(ref[n] ** 2)) for n in ns}
        total = sum(den.values())
        vec = (sum(num.values()) / total) ** 0.5 if total > 0 else sum(num.values()) ** 0.5
        parts = []
        for n in ns:
            share = den[n] / total if total > 0 else 0.0
            own = (num[n] / den[n]) ** 0.5 if den[n] > 0 else 0.0
            parts.append(f"{n} {own:.1e} (share {share:.2f})")
            if share >= 0.1:
                assert own <= tol, f"{label}/{n}: {own} (norm share {share})"
        print(f"[gpu] {label} FP32 {group} vector {vec:.2e}: " + ", ".join(parts), flush=True)
        assert vec <= tol, f"{label}/{group}: {vec}"


DOUBLE_MATERIAL_CASES = [
    # random eps and mu grids on the fine grid (pins the order of the nested 8-cell averages),
    # PML and oblique TF/SF
    ("--3d --size x:13,y:12,z:14 --use-pml --pml-size x:3,y:2,z:3 --use-tfsf --tfsf-size-left x:4,y:4,z:4 "
     "--tfsf-size-right x:4,y:4,z:5 --angle-teta 30 --angle-phi 45 --angle-psi 90 --point-source x:6,y:6,z:7 "
     "--point-source-hz --time-steps 25", ("eps", "mu")),
    # accurate eps sphere and stair mu sphere on the fine grid
    ("--3d --size x:14,y:13,z:12 --use-pml --pml-size x:3,y:3,z:3 --use-tfsf --tfsf-size-left x:4 "
     "--tfsf-size-right x:4 --same-size-tfsf --angle-teta 60 --angle-phi 20 --angle-psi 30 --eps-sphere 3 "
     "--eps-sphere-center x:7,y:6,z:6 --eps-sphere-radius 3 --sphere-accuracy 20 --time-steps 25", ()),
    ("--3d --size x:12,y:14,z:13 --use-tfsf --tfsf-size-left x:3 --tfsf-size-right x:3 --same-size-tfsf "
     "--angle-teta 90 --angle-phi 45 --angle-psi 0 --mu-sphere 2 --mu-sphere-center x:6,y:7,z:6 "
     "--mu-sphere-radius 3 --stair-sphere-approx --current-source x:5,y:6,z:6 --current-source-jz --time-steps 25",
     ()),
    # the 2D schemes: TMz / TEz with kernels of their own, TMx / TEy on the 3D kernels
    ("--2d-tmz --size x:26,y:24 --use-pml --pml-size x:4,y:3 --use-tfsf --tfsf-size-left x:7,y:6 "
     "--tfsf-size-right x:6,y:7 --angle-teta 90 --angle-phi 30 --angle-psi 90 --point-source x:11,y:10 "
     "--point-source-ez --time-steps 40", ("eps", "mu")),
    ("--2d-tez --size x:24,y:26 --use-pml --pml-size x:3,y:4 --use-tfsf --tfsf-size-left x:6,y:7 "
     "--tfsf-size-right x:7,y:6 --angle-teta 90 --angle-phi 30 --angle-psi 0 --eps-sphere 3 "
     "--eps-sphere-center x:12,y:13 --eps-sphere-radius 4 --sphere-accuracy 20 --time-steps 40", ("mu",)),
    ("--2d-tmx --size x:6,y:24,z:22 --use-pml --pml-size y:3,z:3 --use-tfsf --tfsf-size-left y:6,z:6 "
     "--tfsf-size-right y:6,z:6 --angle-teta 35 --angle-phi 90 --angle-psi 0 --time-steps 40", ("eps", "mu")),
    ("--2d-tey --size x:23,y:5,z:25 --use-pml --pml-size x:3,z:4 --point-source x:11,y:0,z:12 "
     "--point-source-hy --time-steps 40", ("eps", "mu")),
    # a 1D line
    ("--1d-ezhy --size x:60 --use-pml --pml-size x:7 --use-tfsf --tfsf-size-left x:15 --tfsf-size-right x:14 "
     "--angle-teta 90 --angle-phi 0 --angle-psi 90 --time-steps 70", ("eps", "mu")),
]


def random_material(cfg, seed, lo=1.0, hi=4.0):
    """A random material grid of the configuration's material shape."""
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=cfg.material_shape)


# Analytic test problems of the reference's test suite (Tests/suite t2-t5: polinom1 in 3D, TMz and
# TEy, polinom2 in 3D, polinom3 in 3D, TMz and TEy, sin1 with eps / mu normed), small sizes, plus a
# norm box, a 1D line and a TF/SF run with a right side only.
_FN = "--use-{k}-border-condition --use-{k}-start-values --use-{k}-right-side --calc-{k}-diff-norm"
TEST_FN_CASES = [
    "--3d --size x:12 --same-size --angle-phi 0 --dx 0.001 --wavelength 0.02 " + _FN.format(k="polinom1") + " --time-steps 21",
    "--2d-tmz --size x:14 --same-size --angle-phi 0 --dx 0.001 --wavelength 0.02 " + _FN.format(k="polinom1") + " --time-steps 21",
    "--2d-tey --size x:14 --same-size --angle-phi 0 --dx 0.001 --wavelength 0.02 " + _FN.format(k="polinom1") + " --time-steps 21",
    "--3d --size x:12 --same-size --angle-phi 0 --dx 0.001 --wavelength 0.02 " + _FN.format(k="polinom2") + " --time-steps 21",
    "--3d --size x:12 --same-size --angle-phi 0 --dx 0.001 --wavelength 0.02 " + _FN.format(k="polinom3") + " --time-steps 21",
    "--2d-tmz --size x:14 --same-size --angle-phi 0 --dx 0.001 --wavelength 0.02 " + _FN.format(k="polinom3") + " --time-steps 21",
    "--3d --size x:10 --same-size --angle-phi 0 --dx 0.001 --wavelength 0.02 --use-sin1-border-condition "
    "--use-sin1-start-values --calc-sin1-diff-norm --eps-normed --mu-normed --courant-factor 149896229 --time-steps 21",
    "--3d --size x:13,y:11,z:12 --angle-phi 0 --dx 0.001 --wavelength 0.02 " + _FN.format(k="polinom2")
    + " --norm-start x:3,z:2 --norm-end y:8 --time-steps 15",
    "--1d-ezhy --size x:40 --angle-teta 90 --angle-phi 0 --angle-psi 90 --dx 0.001 --wavelength 0.02 "
    + _FN.format(k="polinom1") + " --time-steps 30",
    "--3d --size x:16 --same-size --use-tfsf --tfsf-size-left x:4 --tfsf-size-right x:4 --same-size-tfsf "
    "--angle-teta 30 --angle-phi 45 --angle-psi 90 --use-polinom2-right-side --time-steps 20",
]


def run_norms(eng, cfg):
    """Steps `eng` (engine or oracle) through cfg.time_steps one at a time; returns the per-step
    (norm, percent) of every component with an exact solution and the largest |exact|."""
    out = {}
    for _ in range(cfg.time_steps):
        eng.step(1)
        for n in cfg.components:
            try:
                v = eng.diff_norm(n)
            except Exception:
                continue
            out.setdefault(n, []).append(v)
    return {n: np.array(v) for n, v in out.items()}



# Exact solutions exp1 / exp2 / exp3 of the 1D pairs (Tests/suite t7.1 - t7.6, scaled down by 5:
# 120 cells, PML 10, TF/SF border 30 on the left only, a dielectric from 60 on): per pair the line
# axis and the incident angles of its suite script.
EXP_PAIRS = {
    "exhy": ("z", "--angle-teta 0 --angle-phi 0 --angle-psi 90"),
    "exhz": ("y", "--angle-teta 90 --angle-phi 90 --angle-psi 0"),
    "eyhx": ("z", "--angle-teta 0 --angle-phi 90 --angle-psi 90"),
    "eyhz": ("x", "--angle-teta 90 --angle-phi 0 --angle-psi 0"),
    "ezhx": ("y", "--angle-teta 90 --angle-phi 90 --angle-psi 90"),
    "ezhy": ("x", "--angle-teta 90 --angle-phi 0 --angle-psi 90"),
}


def exp_case(pair, kind, size=120, pml=10, tfsf=30, center=90, radius=30, steps=200, dx="0.0001", wavelength="0.004",
             eps=None, norm=None):
    """Command line of a suite t7 run: exp1 in vacuum (eps 1) over the total-field region, exp2
    (reflected) in the scattered-field region left of the border, exp3 (transmitted) inside the
    eps = 4 dielectric."""
    a, angles = EXP_PAIRS[pair]
    if eps is None:
        eps = 1 if kind == "exp1" else 4
    if norm is None:
        b = center - radius
        norm = {"exp1": (tfsf + 1, size - pml - 1), "exp2": (pml + 1, tfsf - 1), "exp3": (b + 1, size - pml - 1)}[kind]
    return (f"--1d-{pair} --size {a}:{size} {angles} --dx {dx} --wavelength {wavelength} --courant-factor 0.5 "
            f"--pml-size {a}:{pml} --use-pml --use-tfsf --tfsf-size-left {a}:{tfsf} --tfsf-size-right {a}:0 "
            f"--eps-sphere {eps} --eps-sphere-center {a}:{center} --eps-sphere-radius {radius} "
            f"--norm-start {a}:{norm[0]} --norm-end {a}:{norm[1]} --calc-{kind}-{pair}-diff-norm --time-steps {steps}")


EXP_CASES = [exp_case(p, k) for p in EXP_PAIRS for k in ("exp1", "exp2", "exp3")]


def run_norms_complex(re_, im_, cfg, mod_norm):
    """Steps the real and imaginary part runs (engine parts or oracles) in lock step; per step the
    complex build's line of every component with an exp solution: (norm re, percent re, norm im,
    percent im, norm of the module, percent).  mod_norm (re_, im_, name) gives the module's."""
    out = {}
    for _ in range(cfg.time_steps):
        re_.step(1)
        im_.step(1)
        for n in cfg.components:
            try:
                r, i = re_.diff_norm(n), im_.diff_norm(n)
            except Exception:
                continue
            m = mod_norm(re_, im_, n)
            out.setdefault(n, []).append((r[0], r[1], i[0], i[1], m[0], m[1]))
    return {n: np.array(v) for n, v in out.items()}
