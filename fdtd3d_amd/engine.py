"""ctypes binding of libyee.so (include/yee.h).

This is the host-side mirror of the reference's Scheme/InternalScheme life cycle for
the Yee path:

    Engine(cfg)          ~ InternalScheme::init + initScheme      (InternalScheme.specific.h:37-67)
    set_material(...)    ~ the material part of Scheme::initGrids (Scheme.inc.h:2962-3165)
    prepare()            ~ initSigmas + coefficient set-up        (Scheme.cpp:355-388)
    step(n)              ~ Scheme::performNStepsForBlock          (Scheme.inc.h:28-473)
    field(name)          ~ reading the newest time layer of a Grid

The binding refuses to run without the native library: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import Dict, Optional

import numpy as np

from .config import Config, COMP_INDEX, MAT_INDEX, YeeConfigC

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libyee.so")

EXPORTED = [
    "yee_last_error", "yee_device_info", "yee_create", "yee_destroy", "yee_set_material",
    "yee_prepare", "yee_step", "yee_sync", "yee_time", "yee_get_field", "yee_get_field_raw", "yee_get_scattered_field", "yee_get_material",
    "yee_get_mask", "yee_get_incident", "yee_get_scalars", "yee_last_step_ms", "yee_profile_steps",
    "yee_get_scattered_field_noalloc", "yee_get_imag_part", "yee_material_size", "yee_set_test_functions",
    "yee_get_diff_norm", "yee_ntff_snapshot", "yee_ntff", "yee_set_exact_exp", "yee_get_diff_norm_complex",
    "yee_set_material_f32", "yee_set_material_sphere", "yee_set_point_source_gaussian",
    "yee_batch_last_error", "yee_batch_create", "yee_batch_destroy", "yee_batch_set_material",
    "yee_batch_set_material_f32", "yee_batch_set_point_source_gaussian", "yee_batch_prepare", "yee_batch_step",
    "yee_batch_sync", "yee_batch_time", "yee_batch_last_step_ms", "yee_batch_tuning", "yee_batch_get_field",
]


# PhysicsConst.h:27-32
MU0 = 4 * 3.14159265358979323846 * 0.0000001
EPS0 = 1 / (MU0 * (299792458.0 * 299792458.0))


class AdjointSpecC(ctypes.Structure):
    """Binary layout of ``struct yee_adjoint_spec``."""
    _fields_ = [("design_lo", ctypes.c_int64 * 3), ("design_hi", ctypes.c_int64 * 3), ("monitor_z", ctypes.c_int64),
                ("monitor_lo", ctypes.c_int64 * 2), ("monitor_hi", ctypes.c_int64 * 2), ("window", ctypes.c_int32),
                ("checkpoint_every", ctypes.c_int32)]


class YeeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"yee error {code}: {msg}")
        self.code = code


_lib = None


def load_library():
    """Loads libyee.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m fdtd3d_amd.build` first")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    i32 = ctypes.c_int32
    L.yee_last_error.restype = ctypes.c_char_p
    L.yee_device_info.argtypes = [ctypes.c_char_p, ctypes.c_int64]
    L.yee_create.argtypes = [vp, ctypes.POINTER(vp)]
    L.yee_destroy.argtypes = [vp]
    L.yee_destroy.restype = None
    L.yee_set_material.argtypes = [vp, i32, vp]
    L.yee_prepare.argtypes = [vp]
    L.yee_step.argtypes = [vp, i32]
    L.yee_sync.argtypes = [vp]
    L.yee_time.argtypes = [vp, ctypes.POINTER(i32)]
    L.yee_get_field.argtypes = [vp, i32, vp]
    L.yee_get_field_raw.argtypes = [vp, i32, vp]
    L.yee_get_scattered_field.argtypes = [vp, i32, vp]
    L.yee_get_scattered_field_noalloc.argtypes = [vp, i32, vp]
    L.yee_get_imag_part.argtypes = [vp, ctypes.POINTER(vp)]
    L.yee_get_material.argtypes = [vp, i32, vp]
    L.yee_material_size.argtypes = [vp, vp]
    L.yee_set_test_functions.argtypes = [vp, i32, i32, i32, i32, vp, vp]
    L.yee_get_diff_norm.argtypes = [vp, i32, vp]
    L.yee_ntff_snapshot.argtypes = [vp]
    L.yee_set_exact_exp.argtypes = [vp, i32, i32, ctypes.c_double, vp, vp]
    L.yee_get_diff_norm_complex.argtypes = [vp, i32, vp]
    L.yee_ntff.argtypes = [vp, ctypes.c_double, vp, i32, vp, vp, vp]
    L.yee_get_mask.argtypes = [vp, i32, i32, vp]
    L.yee_get_incident.argtypes = [vp, vp, vp, ctypes.c_int64]
    L.yee_get_scalars.argtypes = [vp, vp]
    L.yee_last_step_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.yee_profile_steps.argtypes = [vp, i32, vp]
    L.yee_set_material_f32.argtypes = [vp, i32, vp]
    L.yee_set_point_source_gaussian.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.yee_set_material_sphere.argtypes = [vp, i32, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.c_double, i32, i32]
    L.yee_adjoint_setup.argtypes = [vp, ctypes.POINTER(AdjointSpecC)]
    L.yee_forward.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_double)]
    L.yee_backward.argtypes = [vp, ctypes.c_double, vp]
    L.yee_adjoint_timing.argtypes = [vp, vp]
    L.yee_batch_last_error.restype = ctypes.c_char_p
    L.yee_batch_create.argtypes = [vp, i32, ctypes.POINTER(vp)]
    L.yee_batch_destroy.argtypes = [vp]
    L.yee_batch_destroy.restype = None
    L.yee_batch_set_material.argtypes = [vp, i32, i32, vp]
    L.yee_batch_set_material_f32.argtypes = [vp, i32, i32, vp]
    L.yee_batch_set_point_source_gaussian.argtypes = [vp, i32, ctypes.c_double, ctypes.c_double]
    L.yee_batch_prepare.argtypes = [vp]
    L.yee_batch_step.argtypes = [vp, i32]
    L.yee_batch_sync.argtypes = [vp]
    L.yee_batch_time.argtypes = [vp, ctypes.POINTER(i32)]
    L.yee_batch_last_step_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.yee_batch_tuning.argtypes = [vp, vp, vp, i32]
    L.yee_batch_get_field.argtypes = [vp, i32, i32, vp]
    for name in EXPORTED:
        if name not in ("yee_last_error", "yee_destroy", "yee_batch_last_error", "yee_batch_destroy"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise YeeError(rc, load_library().yee_last_error().decode())


def device_info() -> str:
    buf = ctypes.create_string_buffer(256)
    _check(load_library().yee_device_info(buf, 256))
    return buf.value.decode()


class ImagPart:
    """Read access to the imaginary part of a complex-field Engine (a handle it owns)."""

    def __init__(self, owner: "Engine", h):
        self._owner = owner   # keeps the owning handle alive
        self.h = h
        self.shape = owner.shape
        self.dtype = owner.dtype

    def field(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        _check(load_library().yee_get_field(self.h, COMP_INDEX[name], out.ctypes.data))
        return out

    def field_raw(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(load_library().yee_get_field_raw(self.h, COMP_INDEX[name], out.ctypes.data))
        return out

    def scattered(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        _check(load_library().yee_get_scattered_field(self.h, COMP_INDEX[name], out.ctypes.data))
        return out

    def incident(self, n: int):
        e = np.empty(n, dtype=np.float64)
        hh = np.empty(n, dtype=np.float64)
        _check(load_library().yee_get_incident(self.h, e.ctypes.data, hh.ctypes.data, int(n)))
        return e, hh


class Engine:
    """One simulation on the GPU (a yee_handle)."""

    def __init__(self, cfg: Config):
        L = load_library()
        self.cfg = cfg
        self._c = cfg.to_c()
        h = ctypes.c_void_p()
        _check(L.yee_create(ctypes.byref(self._c), ctypes.byref(h)))
        self.h = h
        self.shape = cfg.shape
        self.dtype = np.float32 if cfg.precision == 4 else np.float64
        # material grids: the field grid's shape, or the fine grid of double material precision
        ms = (ctypes.c_int64 * 3)()
        _check(L.yee_material_size(h, ms))
        self.material_shape = tuple(int(ms[a]) for a in range(3) if cfg.dim == 3 or a in cfg.active_axes)
        if cfg.start_fn or cfg.border_fn or cfg.rhs_fn or cfg.exact_fn:
            # analytic test problem of the reference's --use-<kind>-* / --calc-<kind>-diff-norm
            ns = (ctypes.c_int64 * 3)(*[int(v) for v in cfg.norm_start])
            ne = (ctypes.c_int64 * 3)(*[int(v) for v in cfg.norm_end])
            _check(L.yee_set_test_functions(h, cfg.start_fn, cfg.border_fn, cfg.rhs_fn, cfg.exact_fn, ns, ne))
        if cfg.exp_fn:
            # --calc-exp<n>-<pair>-diff-norm
            ns = (ctypes.c_int64 * 3)(*[int(v) for v in cfg.norm_start])
            ne = (ctypes.c_int64 * 3)(*[int(v) for v in cfg.norm_end])
            _check(L.yee_set_exact_exp(h, cfg.exp_fn, cfg.exp_pair, cfg.exp_border, ns, ne))

    def close(self) -> None:
        if getattr(self, "h", None):
            load_library().yee_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_material(self, name: str, values: np.ndarray) -> None:
        if int(np.size(values)) != int(np.prod(self.material_shape)):
            raise ValueError(f"material {name}: expected {self.material_shape}, got {np.shape(values)}")
        if isinstance(values, np.ndarray) and values.dtype == np.float32:
            arr = np.ascontiguousarray(values)
            _check(load_library().yee_set_material_f32(self.h, MAT_INDEX[name], arr.ctypes.data))
            return
        arr = np.ascontiguousarray(values, dtype=np.float64)
        _check(load_library().yee_set_material(self.h, MAT_INDEX[name], arr.ctypes.data))

    def set_material_sphere(self, name: str, value: float, outside: float, center, radius: float,
                            accuracy: int = 100, stair: bool = False) -> None:
        """Sphere material as Scheme::initGrids builds it (center in cell indices)."""
        c = (ctypes.c_int64 * 3)(*[int(v) for v in (list(center) + [0, 0, 0])[:3]])
        _check(load_library().yee_set_material_sphere(self.h, MAT_INDEX[name], float(value), float(outside), c,
                                                      float(radius), int(accuracy), 1 if stair else 0))

    def set_point_source_gaussian(self, t0: float, sigma: float) -> None:
        """Gaussian envelope exp(-((ts - t0) / sigma)^2) on the hard point source (steps)."""
        _check(load_library().yee_set_point_source_gaussian(self.h, float(t0), float(sigma)))

    def prepare(self) -> None:
        # --eps-normed / --mu-normed replace the grids by 1 / eps0, 1 / mu0 (Scheme.inc.h:3016-3025, 3066-3075)
        for on, name, v in ((self.cfg.eps_normed, "eps", 1.0 / EPS0), (self.cfg.mu_normed, "mu", 1.0 / MU0)):
            if on:
                self.set_material(name, np.full(self.material_shape, v))
        _check(load_library().yee_prepare(self.h))

    def diff_norm(self, name: str):
        """(norm, percent, sum of squares, max |exact|) of a component against the exact solution
        at the last step (the reference's "DIFF NORM" line)."""
        out = (ctypes.c_double * 4)()
        _check(load_library().yee_get_diff_norm(self.h, COMP_INDEX[name], out))
        return tuple(out)

    def diff_norm_complex(self, name: str):
        """The complex build's diff norm of a component at the last step: (norm re, norm im,
        norm of the module, their percentages, max |Re|, max |Im|, max |exact|)."""
        out = (ctypes.c_double * 9)()
        _check(load_library().yee_get_diff_norm_complex(self.h, COMP_INDEX[name], out))
        return tuple(out)

    def ntff_snapshot(self) -> None:
        """Keeps H of both parts for the near-to-far-field transform one step later."""
        _check(load_library().yee_ntff_snapshot(self.h))

    def ntff(self, teta: float, phis, left, right) -> np.ndarray:
        """Pointing_scat / Pointing_inc at elevation teta and each azimuth of phis (radians) over
        the box [left, right] (Scheme::saveNTFF, Scheme.inc.h:4535-4640); needs ntff_snapshot
        one step before."""
        phi = np.ascontiguousarray(np.atleast_1d(np.asarray(phis, dtype=np.float64)))
        lo = (ctypes.c_int64 * 3)(*[int(v) for v in left])
        hi = (ctypes.c_int64 * 3)(*[int(v) for v in right])
        out = np.zeros(phi.size)
        _check(load_library().yee_ntff(self.h, float(teta), phi.ctypes.data, int(phi.size), lo, hi, out.ctypes.data))
        return out

    def step(self, n: int) -> None:
        _check(load_library().yee_step(self.h, int(n)))

    def sync(self) -> None:
        _check(load_library().yee_sync(self.h))

    @property
    def time(self) -> int:
        t = ctypes.c_int32()
        _check(load_library().yee_time(self.h, ctypes.byref(t)))
        return t.value

    def last_step_ms(self) -> float:
        ms = ctypes.c_double()
        _check(load_library().yee_last_step_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def profile_steps(self, n: int) -> dict:
        """Runs n steps with per-kernel HIP events; average ms per launch of each kernel."""
        out = np.zeros(5, dtype=np.float64)
        _check(load_library().yee_profile_steps(self.h, int(n), out.ctypes.data))
        return {"incident_ms": float(out[0]), "e_half_ms": float(out[1]), "h_half_ms": float(out[2]),
                "fused_ms": float(out[3]), "step_ms": float(out[4])}

    def adjoint_setup(self, design_lo, design_hi, monitor_z, monitor_lo, monitor_hi, window, checkpoint_every=0):
        sp = AdjointSpecC()
        for a in range(3):
            sp.design_lo[a], sp.design_hi[a] = int(design_lo[a]), int(design_hi[a])
        sp.monitor_z = int(monitor_z)
        for a in range(2):
            sp.monitor_lo[a], sp.monitor_hi[a] = int(monitor_lo[a]), int(monitor_hi[a])
        sp.window = int(window)
        sp.checkpoint_every = int(checkpoint_every)
        self._design_shape = tuple(int(design_hi[a]) - int(design_lo[a]) for a in range(3))
        _check(load_library().yee_adjoint_setup(self.h, ctypes.byref(sp)))

    def forward(self, design_eps: np.ndarray) -> float:
        arr = np.ascontiguousarray(design_eps, dtype=np.float64)
        if arr.shape != self._design_shape:
            raise ValueError(f"design eps: expected {self._design_shape}, got {arr.shape}")
        J = ctypes.c_double()
        _check(load_library().yee_forward(self.h, arr.ctypes.data, ctypes.byref(J)))
        return J.value

    def backward(self, dJ: float = 1.0) -> np.ndarray:
        out = np.empty(self._design_shape, dtype=np.float64)
        _check(load_library().yee_backward(self.h, float(dJ), out.ctypes.data))
        return out

    def adjoint_timing(self):
        out = np.zeros(2, dtype=np.float64)
        _check(load_library().yee_adjoint_timing(self.h, out.ctypes.data))
        return {"forward_ms": float(out[0]), "backward_ms": float(out[1])}

    def field(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        _check(load_library().yee_get_field(self.h, COMP_INDEX[name], out.ctypes.data))
        return out

    def imag(self) -> "ImagPart":
        """The imaginary part of a complex-field run (the real part is this handle)."""
        h = ctypes.c_void_p()
        _check(load_library().yee_get_imag_part(self.h, ctypes.byref(h)))
        return ImagPart(self, h)

    def scattered(self, name: str) -> np.ndarray:
        """Scattered field: the newest field minus the incident wave inside the TF/SF box."""
        out = np.empty(self.shape, dtype=np.float64)
        _check(load_library().yee_get_scattered_field(self.h, COMP_INDEX[name], out.ctypes.data))
        return out

    def material(self, name: str) -> np.ndarray:
        """Material grid as the engine holds it (relative eps / mu; omega_p, gamma in rad/s)."""
        out = np.empty(self.material_shape, dtype=np.float64)
        _check(load_library().yee_get_material(self.h, MAT_INDEX[name], out.ctypes.data))
        return out

    def field_raw(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(load_library().yee_get_field_raw(self.h, COMP_INDEX[name], out.ctypes.data))
        return out

    def tfsf_mask(self, comp: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.uint8)
        _check(load_library().yee_get_mask(self.h, 0, COMP_INDEX[comp], out.ctypes.data))
        return out

    def special_mask(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.uint8)
        _check(load_library().yee_get_mask(self.h, 1, 0, out.ctypes.data))
        return out

    def incident(self, length: Optional[int] = None):
        n = 500 * self.cfg.size[0] if length is None else length
        e = np.empty(n, dtype=np.float64)
        h = np.empty(n, dtype=np.float64)
        _check(load_library().yee_get_incident(self.h, e.ctypes.data, h.ctypes.data, n))
        return e, h

    def scalars(self) -> np.ndarray:
        out = np.empty(3, dtype=np.float64)
        _check(load_library().yee_get_scalars(self.h, out.ctypes.data))
        return out


def _bcheck(rc: int) -> None:
    if rc != 0:
        raise YeeError(rc, load_library().yee_batch_last_error().decode())


class Batch:
    """Batched TMz sweep (yee_batch): independent TMz simulations advanced together."""

    def __init__(self, cfgs):
        L = load_library()
        self.cfgs = list(cfgs)
        arr = (YeeConfigC * len(self.cfgs))(*[c.to_c() for c in self.cfgs])
        self._arr = arr
        h = ctypes.c_void_p()
        _bcheck(L.yee_batch_create(arr, len(self.cfgs), ctypes.byref(h)))
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            load_library().yee_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_material(self, sim: int, name: str, values: np.ndarray) -> None:
        c = self.cfgs[sim]
        if int(np.size(values)) != int(np.prod(c.shape)):
            raise ValueError(f"material {name}: expected {c.shape}, got {np.shape(values)}")
        if isinstance(values, np.ndarray) and values.dtype == np.float32:
            arr = np.ascontiguousarray(values)
            _bcheck(load_library().yee_batch_set_material_f32(self.h, sim, MAT_INDEX[name], arr.ctypes.data))
            return
        arr = np.ascontiguousarray(values, dtype=np.float64)
        _bcheck(load_library().yee_batch_set_material(self.h, sim, MAT_INDEX[name], arr.ctypes.data))

    def set_point_source_gaussian(self, sim: int, t0: float, sigma: float) -> None:
        _bcheck(load_library().yee_batch_set_point_source_gaussian(self.h, sim, float(t0), float(sigma)))

    def prepare(self) -> None:
        _bcheck(load_library().yee_batch_prepare(self.h))

    def step(self, n: int) -> None:
        _bcheck(load_library().yee_batch_step(self.h, int(n)))

    def sync(self) -> None:
        _bcheck(load_library().yee_batch_sync(self.h))

    @property
    def time(self) -> int:
        t = ctypes.c_int32()
        _bcheck(load_library().yee_batch_time(self.h, ctypes.byref(t)))
        return t.value

    def last_step_ms(self) -> float:
        ms = ctypes.c_double()
        _bcheck(load_library().yee_batch_last_step_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def tuning(self, ncand: int = 12):
        out = np.zeros(3, dtype=np.int32)
        ms = np.zeros(ncand, dtype=np.float64)
        _bcheck(load_library().yee_batch_tuning(self.h, out.ctypes.data, ms.ctypes.data, ncand))
        return {"tile_rows": int(out[0]), "workgroups": int(out[1]), "groups": int(out[2]),
                "candidate_ms": [float(v) for v in ms]}

    def field(self, sim: int, name: str) -> np.ndarray:
        c = self.cfgs[sim]
        out = np.empty(c.shape, dtype=np.float64)
        _bcheck(load_library().yee_batch_get_field(self.h, sim, COMP_INDEX[name], out.ctypes.data))
        return out


def run_engine(cfg: Config, materials: Optional[Dict[str, np.ndarray]] = None, steps: Optional[int] = None) -> Engine:
    eng = Engine(cfg)
    for name, arr in (materials or {}).items():
        eng.set_material(name, arr)
    eng.prepare()
    eng.step(cfg.time_steps if steps is None else steps)
    eng.sync()
    return eng


def ntff_angles(cfg: Config) -> np.ndarray:
    """Azimuths (radians) of one NTFF diagram (Scheme::saveNTFF, Scheme.inc.h:4561-4605): the
    incident wave's phi with --ntff-reverse, else 0, step, 2 step, ... while <= 2 pi + pi / 180,
    step = pi * --ntff-step-angle / 180, added up one by one as the reference's loop does."""
    if cfg.ntff_reverse:
        return np.array([cfg.angle_deg[1] * math.pi / 180.0])
    end, step = 2 * math.pi + math.pi / 180, math.pi * cfg.ntff_step_angle / 180
    if not step > 0:
        raise ValueError("--ntff-step-angle must be positive")
    out, a = [], 0.0
    while a <= end:
        out.append(a)
        a += step
    return np.array(out)


def ntff_diagrams(eng: "Engine", cfg: Config):
    """The diagrams Scheme::saveNTFF writes at the engine's current step, one per box step s of
    --ntff-diff: (s, left, right, phis, values), box [ntff_size + s, size - ntff_size + 1 - s],
    elevation the incident wave's teta.  The engine must have taken ntff_snapshot one step before."""
    teta = cfg.angle_deg[0] * math.pi / 180.0
    phis = ntff_angles(cfg)
    out = []
    for s in range(cfg.ntff_diff):
        left = [cfg.ntff_size[a] + s for a in range(3)]
        right = [cfg.size[a] - cfg.ntff_size[a] + 1 - s for a in range(3)]
        out.append((s, left, right, phis, eng.ntff(teta, phis, left, right)))
    return out

