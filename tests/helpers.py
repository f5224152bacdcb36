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


def golden_names():
    return sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


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
