"""ctypes binding of the C oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; the product never does.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import subprocess

import numpy as np

from fdtd3d_amd.config import Config, COMP_INDEX, MAT_INDEX

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [vp]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_set_material.argtypes = [vp, ctypes.c_int32, vp]
        L.oracle_prepare.argtypes = [vp]
        L.oracle_step.argtypes = [vp, ctypes.c_int32]
        L.oracle_get_field.argtypes = [vp, ctypes.c_int32, vp]
        L.oracle_get_material.argtypes = [vp, ctypes.c_int32, vp]
        L.oracle_tfsf_mask.argtypes = [vp, ctypes.c_int32, vp]
        L.oracle_get_incident.argtypes = [vp, vp, vp, ctypes.c_int64]
        L.oracle_get_scattered.argtypes = [vp, ctypes.c_int32, vp]
        L.oracle_get_scalars.argtypes = [vp, vp]
        L.oracle_incident_len.restype = ctypes.c_int64
        L.oracle_incident_len.argtypes = [vp]
        L.oracle_set_source_phase.argtypes = [vp, ctypes.c_double]
        L.oracle_set_source_kind.argtypes = [vp, ctypes.c_int]
        L.oracle_set_source_gaussian.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.oracle_set_test_functions.argtypes = [vp] + [ctypes.c_int] * 4 + [vp, vp]
        L.oracle_get_diff_norm.argtypes = [vp, ctypes.c_int32, vp]
        L.oracle_ntff.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, vp, vp, vp]
        L.oracle_set_exact_exp.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp, vp]
        L.oracle_exp_mod_norm.argtypes = [vp, vp, ctypes.c_int32, vp]
        L.oracle_scale_fields.argtypes = [vp, ctypes.c_double]
        L.oracle_scale_fields.restype = None
        L.oracle_adjoint.argtypes = [vp, ctypes.POINTER(OracleAdjSpec), ctypes.c_int32,
                                     ctypes.POINTER(ctypes.c_double), vp]
        d = ctypes.c_double
        L.oracle_sphere_accurate_3d.restype = d
        L.oracle_sphere_accurate_3d.argtypes = [d] * 9 + [ctypes.c_int]
        L.oracle_sphere_accurate_2d.restype = d
        L.oracle_sphere_accurate_2d.argtypes = [d] * 7 + [ctypes.c_int]
        L.oracle_sphere_accurate_1d.restype = d
        L.oracle_sphere_accurate_1d.argtypes = [d] * 5
        L.oracle_sphere_stair_1d.restype = d
        L.oracle_sphere_stair_1d.argtypes = [d] * 5
        L.oracle_sphere_stair_3d.restype = d
        L.oracle_sphere_stair_3d.argtypes = [d] * 9
        L.oracle_wave_number.restype = d
        L.oracle_wave_number.argtypes = [d] * 6
        L.oracle_wave_number_general.restype = d
        L.oracle_wave_number_general.argtypes = [d] * 6
        _lib = L
    return _lib


class OracleAdjSpec(ctypes.Structure):
    _fields_ = [("dlo", ctypes.c_int64 * 3), ("dhi", ctypes.c_int64 * 3), ("mz", ctypes.c_int64),
                ("mlo", ctypes.c_int64 * 2), ("mhi", ctypes.c_int64 * 2), ("window", ctypes.c_int32)]


def oracle_adjoint(cfg: Config, eps: np.ndarray, design_lo, design_hi, monitor_z: int, monitor_lo, monitor_hi,
                   window: int, steps: int):
    """FP64 CPU adjoint of the reference-form update: (J, dJ/deps over the design box)."""
    o = Oracle(cfg)
    o.set_material("eps", eps)
    o.prepare()
    sp = OracleAdjSpec()
    for a in range(3):
        sp.dlo[a], sp.dhi[a] = int(design_lo[a]), int(design_hi[a])
    sp.mz = int(monitor_z)
    for a in range(2):
        sp.mlo[a], sp.mhi[a] = int(monitor_lo[a]), int(monitor_hi[a])
    sp.window = int(window)
    shape = tuple(int(design_hi[a]) - int(design_lo[a]) for a in range(3))
    grad = np.zeros(shape, dtype=np.float64)
    J = ctypes.c_double()
    rc = lib().oracle_adjoint(o.h, ctypes.byref(sp), int(steps), ctypes.byref(J), grad.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle_adjoint failed ({rc})")
    o.close()
    return J.value, grad


def oracle_objective(cfg: Config, eps: np.ndarray, monitor_z: int, monitor_lo, monitor_hi, window: int, steps: int):
    """J of the forward oracle alone (for finite differences)."""
    o = Oracle(cfg)
    o.set_material("eps", eps)
    o.prepare()
    J = 0.0
    for t in range(steps):
        o.step(1)
        if t >= steps - window:
            ez = o.field("Ez")[monitor_lo[0]:monitor_hi[0], monitor_lo[1]:monitor_hi[1], monitor_z]
            J += float(np.sum(ez * ez))
    o.close()
    return J


def sphere_map(cfg: Config, value: float, outside: float, center, radius: int) -> np.ndarray:
    """Material map of a sphere as Scheme::initGrids builds it (Scheme.inc.h:2983-3012); with
    --use-double-material-precision on the fine grid, centre 2 c + 0.5 and radius 2 r."""
    L = lib()
    shape = cfg.material_shape
    mod = 2.0 if cfg.double_material else 1.0
    out = np.empty(shape, dtype=np.float64)
    cx, cy, cz = (float(center[0]) * mod + 0.5, float(center[1]) * mod + 0.5, float(center[2]) * mod + 0.5)
    r = float(radius) * mod
    it = np.ndindex(*shape)
    for idx in it:
        if cfg.dim == 3:
            i, j, k = idx
            mx, my, mz = i + 0.5, j + 0.5, k + 0.5
            if cfg.stair_sphere:
                out[idx] = L.oracle_sphere_stair_3d(mx, my, mz, cx, cy, cz, r, value, outside)
            else:
                out[idx] = L.oracle_sphere_accurate_3d(mx, my, mz, cx, cy, cz, r, value, outside, cfg.sphere_accuracy)
        elif cfg.dim == 1:
            # the segment on the line's one coordinate
            c = (cx, cy, cz)[cfg.active_axes[0]]
            m = idx[0] + 0.5
            out[idx] = (L.oracle_sphere_stair_1d(m, c, r, value, outside) if cfg.stair_sphere
                        else L.oracle_sphere_accurate_1d(m, c, r, value, outside))
        else:
            # the two axes of the 2D scheme (the sphere centre is given in 3D coordinates)
            ca, cb = [c for a, c in enumerate((cx, cy, cz)) if a != cfg.missing_axis]
            i, j = idx
            out[idx] = L.oracle_sphere_accurate_2d(i + 0.5, j + 0.5, ca, cb, r, value, outside, cfg.sphere_accuracy)
    return out


class Oracle:
    """FP64 CPU restatement of the reference path; same life cycle as the engine handle."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self._c = cfg.to_c()
        self.h = lib().oracle_create(ctypes.byref(self._c))
        if not self.h:
            raise ValueError("oracle_create rejected the configuration")
        self.shape = cfg.shape
        self.material_shape = cfg.material_shape
        if cfg.start_fn or cfg.border_fn or cfg.rhs_fn or cfg.exact_fn:
            ns = (ctypes.c_int64 * 3)(*[int(v) for v in cfg.norm_start])
            ne = (ctypes.c_int64 * 3)(*[int(v) for v in cfg.norm_end])
            if lib().oracle_set_test_functions(self.h, cfg.start_fn, cfg.border_fn, cfg.rhs_fn, cfg.exact_fn, ns, ne):
                raise ValueError("oracle_set_test_functions rejected the configuration")
        if cfg.exp_fn:
            ns = (ctypes.c_int64 * 3)(*[int(v) for v in cfg.norm_start])
            ne = (ctypes.c_int64 * 3)(*[int(v) for v in cfg.norm_end])
            if lib().oracle_set_exact_exp(self.h, cfg.exp_fn, cfg.exp_pair, cfg.exp_border, ns, ne):
                raise ValueError("oracle_set_exact_exp rejected the configuration")

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_source_phase(self, phase: float) -> None:
        lib().oracle_set_source_phase(self.h, float(phase))

    def set_source_kind(self, kind: int) -> None:
        """1 / 2: the real / imaginary part of the complex build's sources; 0: the real build."""
        if lib().oracle_set_source_kind(self.h, int(kind)) != 0:
            raise ValueError(f"source kind {kind}")

    def set_source_gaussian(self, t0: float, sigma: float) -> None:
        lib().oracle_set_source_gaussian(self.h, float(t0), float(sigma))

    def set_material(self, name: str, values: np.ndarray) -> None:
        arr = np.ascontiguousarray(values, dtype=np.float64).reshape(self.material_shape)
        rc = lib().oracle_set_material(self.h, MAT_INDEX[name], arr.ctypes.data)
        if rc:
            raise ValueError(name)

    def prepare(self) -> None:
        # --eps-normed / --mu-normed (Scheme.inc.h:3016-3025, 3066-3075)
        mu0 = 4 * 3.14159265358979323846 * 0.0000001
        eps0 = 1 / (mu0 * (299792458.0 * 299792458.0))
        for on, name, v in ((self.cfg.eps_normed, "eps", 1.0 / eps0), (self.cfg.mu_normed, "mu", 1.0 / mu0)):
            if on:
                self.set_material(name, np.full(self.material_shape, v))
        lib().oracle_prepare(self.h)

    def diff_norm(self, name: str):
        """(norm, percent, sum of squares, max |exact|) at the last step."""
        out = (ctypes.c_double * 4)()
        if lib().oracle_get_diff_norm(self.h, COMP_INDEX[name], out):
            raise KeyError(name)
        return tuple(out)

    def scale_fields(self, s: float) -> None:
        """Multiplies every E / H value of every time layer by s."""
        lib().oracle_scale_fields(self.h, float(s))

    def step(self, n: int) -> None:
        if lib().oracle_step(self.h, int(n)):
            raise RuntimeError("oracle_step")

    def field(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        if lib().oracle_get_field(self.h, COMP_INDEX[name], out.ctypes.data):
            raise KeyError(name)
        return out

    def material(self, which: int) -> np.ndarray:
        out = np.empty(self.material_shape, dtype=np.float64)
        if lib().oracle_get_material(self.h, which, out.ctypes.data):
            raise KeyError(which)
        return out

    def tfsf_mask(self, comp: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.uint8)
        if lib().oracle_tfsf_mask(self.h, COMP_INDEX[comp], out.ctypes.data):
            raise KeyError(comp)
        return out

    def scattered(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        if lib().oracle_get_scattered(self.h, COMP_INDEX[name], out.ctypes.data):
            raise KeyError(name)
        return out

    def incident(self):
        n = lib().oracle_incident_len(self.h)
        e = np.empty(n, dtype=np.float64)
        h = np.empty(n, dtype=np.float64)
        lib().oracle_get_incident(self.h, e.ctypes.data, h.ctypes.data, n)
        return e, h

    def scalars(self):
        out = np.empty(3, dtype=np.float64)
        lib().oracle_get_scalars(self.h, out.ctypes.data)
        return out


def run_oracle(cfg: Config, materials=None, steps=None, source_kind: int = 0) -> Oracle:
    o = Oracle(cfg)
    if source_kind:
        o.set_source_kind(source_kind)
    for name, arr in (materials or {}).items():
        o.set_material(name, arr)
    o.prepare()
    o.step(cfg.time_steps if steps is None else steps)
    return o


def run_oracle_complex(cfg: Config, materials=None, steps=None):
    """The reference's complex-field build: its real and imaginary parts, each a real run driven
    by its part of the complex sources (the update multiplies fields by real coefficients only)."""
    real = dataclasses.replace(cfg, complex_fields=False)
    return (run_oracle(real, materials, steps, source_kind=1), run_oracle(real, materials, steps, source_kind=2))


def read_ref_output(path: str):
    """Records written by oracle/ref_driver.cpp: [name:16][int64 n][n doubles]."""
    out = {}
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0
    while pos < len(data):
        name = data[pos:pos + 16].split(b"\0", 1)[0].decode()
        n = int(np.frombuffer(data[pos + 16:pos + 24], dtype=np.int64)[0])
        pos += 24
        out[name] = np.frombuffer(data[pos:pos + 8 * n], dtype=np.float64).copy()
        pos += 8 * n
    return out


def ref_driver_path(kind: str = "f64") -> str:
    return os.path.join(REF_DIR, "refdrv_" + kind)


def run_ref_driver(args, out_path: str, kind: str = "f64", timeout: float = 600):
    env = dict(os.environ)
    env["YEE_REF_OUT"] = out_path
    res = subprocess.run([ref_driver_path(kind)] + list(args), env=env, capture_output=True,
                         text=True, timeout=timeout)
    if res.returncode != 0:
        raise RuntimeError(f"ref driver failed ({res.returncode}): {res.stderr[-2000:]}")
    secs = None
    for line in res.stdout.splitlines():
        if line.startswith("Total time ="):
            secs = float(line.split("=")[1].split()[0])
    return read_ref_output(out_path), secs


def oracle_exp_mod_norm(re: "Oracle", im: "Oracle", name: str):
    """(norm, percent) of the module of the complex build's exp exact solution at the last step,
    from the real and imaginary part oracles."""
    out = (ctypes.c_double * 2)()
    if lib().oracle_exp_mod_norm(re.h, im.h, COMP_INDEX[name], out):
        raise ValueError("oracle_exp_mod_norm rejected the call")
    return out[0], out[1]


def oracle_ntff(re: "Oracle", im: "Oracle", teta: float, phi: float, left, right):
    """(Pointing_scat / Pointing_inc, Pointing_scat) of the complex build at (teta, phi) from the
    real / imaginary part oracles (run_oracle_complex) over the box faces left, right."""
    lo = (ctypes.c_int64 * 3)(*[int(v) for v in left])
    hi = (ctypes.c_int64 * 3)(*[int(v) for v in right])
    out = (ctypes.c_double * 2)()
    if lib().oracle_ntff(re.h, im.h, float(teta), float(phi), lo, hi, out):
        raise ValueError("oracle_ntff rejected the box")
    return out[0], out[1]

