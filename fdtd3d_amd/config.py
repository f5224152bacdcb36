"""Run configuration of the Yee engine: the ctypes mirror of ``yee_config``
(include/yee.h) and a parser for the reference's option table.

The option names, their argument syntax (``x:<int>,y:<int>,z:<int>``) and the
``--same-size*`` copy rules follow the reference's Settings table and parser
(Source/Settings/Settings.inc.h:44-343, Source/Settings/Settings.cpp:133-369).
Floating-point options are parsed the way the reference parses them: with
``std::stof`` (Source/Helpers/PAssert.h:169-170), i.e. rounded to single
precision and then widened, so ``--dx 0.0005`` yields the same double as in
the reference build.
"""
from __future__ import annotations

import ctypes
import re
import struct
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

SCHEME_EXHY = 1
SCHEME_EXHZ = 2
SCHEME_EYHX = 3
SCHEME_EYHZ = 4
SCHEME_EZHX = 5
SCHEME_EZHY = 6
SCHEME_TEX = 7
SCHEME_TEY = 8
SCHEME_TEZ = 9
SCHEME_TMX = 10
SCHEME_TMY = 11
SCHEME_TMZ = 12
SCHEME_DIM3 = 13
# the option of each 2D scheme (Settings.inc.h) and the axis that scheme lacks
# 1D schemes: (scheme, the axis of the grid, its components) (main.cpp:629-800)
SCHEMES_1D = {"--1d-exhy": (SCHEME_EXHY, 2, ("Ex", "Hy")), "--1d-exhz": (SCHEME_EXHZ, 1, ("Ex", "Hz")),
              "--1d-eyhx": (SCHEME_EYHX, 2, ("Ey", "Hx")), "--1d-eyhz": (SCHEME_EYHZ, 0, ("Ey", "Hz")),
              "--1d-ezhx": (SCHEME_EZHX, 1, ("Ez", "Hx")), "--1d-ezhy": (SCHEME_EZHY, 0, ("Ez", "Hy"))}
SCHEMES_2D = {"--2d-tex": (SCHEME_TEX, 0), "--2d-tey": (SCHEME_TEY, 1), "--2d-tez": (SCHEME_TEZ, 2),
              "--2d-tmx": (SCHEME_TMX, 0), "--2d-tmy": (SCHEME_TMY, 1), "--2d-tmz": (SCHEME_TMZ, 2)}

USE_PML = 0x1
USE_TFSF = 0x2
USE_METAMATERIALS = 0x4
COMPLEX_FIELDS = 0x8
DOUBLE_MATERIAL_PRECISION = 0x10
# analytic test problems (include/yee.h YEE_FN_*), in the priority order of Scheme::initCallBacks
TEST_FUNCTIONS = {"polinom1": 1, "polinom2": 2, "polinom3": 3, "sin1": 4}
# exact solutions of a 1D pair (--calc-exp<n>-<pair>-diff-norm): YEE_FN_EXP1 .. EXP3
EXP_FUNCTIONS = {"exp1": 5, "exp2": 6, "exp3": 7}

COMPONENTS = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz",
              "Dx", "Dy", "Dz", "Bx", "By", "Bz",
              "D1x", "D1y", "D1z", "B1x", "B1y", "B1z"]
COMP_INDEX = {name: i for i, name in enumerate(COMPONENTS)}

MATERIALS = ["eps", "mu", "omegape", "omegapm", "gammae", "gammam"]
MAT_INDEX = {name: i for i, name in enumerate(MATERIALS)}


class YeeConfigC(ctypes.Structure):
    """Binary layout of ``struct yee_config``."""
    _fields_ = [
        ("scheme", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("time_steps", ctypes.c_int32),
        ("size", ctypes.c_int64 * 3),
        ("pml", ctypes.c_int64 * 3),
        ("tfsf_left", ctypes.c_int64 * 3),
        ("tfsf_right", ctypes.c_int64 * 3),
        ("dx", ctypes.c_double),
        ("wavelength", ctypes.c_double),
        ("courant", ctypes.c_double),
        ("angle_deg", ctypes.c_double * 3),
        ("flags", ctypes.c_uint32),
        ("point_source", ctypes.c_uint32),
        ("point_source_pos", ctypes.c_int64 * 3),
        ("current_source", ctypes.c_uint32),
        ("current_source_pos", ctypes.c_int64 * 3),
    ]


def stof(text) -> float:
    """Value of ``std::stof(text)`` widened to double (single-precision rounding)."""
    return struct.unpack("f", struct.pack("f", float(text)))[0]


@dataclass
class Config:
    """Python view of ``yee_config`` plus the material options of the reference CLI."""
    scheme: int = SCHEME_DIM3
    precision: int = 8
    time_steps: int = 100                     # --time-steps (reference default 100)
    size: List[int] = field(default_factory=lambda: [100, 100, 100])
    pml: List[int] = field(default_factory=lambda: [2, 2, 2])
    tfsf_left: List[int] = field(default_factory=lambda: [4, 4, 4])
    tfsf_right: List[int] = field(default_factory=lambda: [4, 4, 4])
    # defaults are the table's FPValue literals (exact doubles); values given on the
    # command line go through stof
    dx: float = 0.0005
    wavelength: float = 0.02
    courant: float = 0.5
    angle_deg: List[float] = field(default_factory=lambda: [90.0, 0.0, 90.0])
    use_pml: bool = False
    use_tfsf: bool = False
    use_metamaterials: bool = False
    # complex field values (the reference's -DCOMPLEX_FIELD_VALUES build; no command-line flag)
    complex_fields: bool = False
    # --use-double-material-precision: material grids of twice the field grid's extent per axis
    double_material: bool = False
    # analytic test problems (Settings.inc.h:134-178): kinds (TEST_FUNCTIONS values, 0 = none) of
    # the start values, border values, right side and exact solution; the norm box
    start_fn: int = 0
    border_fn: int = 0
    rhs_fn: int = 0
    exact_fn: int = 0
    norm_start: List[int] = field(default_factory=lambda: [0, 0, 0])
    norm_end: List[int] = field(default_factory=lambda: [0, 0, 0])
    # --eps-normed / --mu-normed: eps = 1 / eps0 / mu = 1 / mu0 everywhere (Scheme.inc.h:3016-3025, 3066-3075)
    eps_normed: bool = False
    mu_normed: bool = False
    # near-to-far-field transform (Settings.inc.h:85-88, 124, 181-187): --use-ntff needs complex
    # field values (Scheme::Pointing_scat asserts the complex build, Scheme.inc.h:4096-4098)
    # --calc-exp<n>-<pair>-diff-norm: exp kind (EXP_FUNCTIONS value, 0 = none) and the pair's
    # 1D scheme value (SCHEME_EXHY .. SCHEME_EZHY)
    exp_fn: int = 0
    exp_pair: int = 0
    use_ntff: bool = False
    ntff_size: List[int] = field(default_factory=lambda: [3, 3, 3])
    ntff_diff: int = 1
    ntff_step_angle: float = 10.0
    ntff_filename: str = "ntff-res"
    ntff_to_stdout: bool = False
    ntff_reverse: bool = False
    interm_ntff_step: int = 100
    point_source: List[str] = field(default_factory=list)   # component names, e.g. ["Ez"]
    point_source_pos: List[int] = field(default_factory=lambda: [0, 0, 0])
    # --current-source-j<x|y|z> / -m<x|y|z> as component names ("Ex" for Jx, "Hx" for Mx), --current-source
    current_source: List[str] = field(default_factory=list)
    current_source_pos: List[int] = field(default_factory=lambda: [0, 0, 0])
    # material options (Settings.inc.h:236-266)
    eps_file: str = ""
    mu_file: str = ""
    omegape_file: str = ""
    omegapm_file: str = ""
    gammae_file: str = ""
    gammam_file: str = ""
    eps_sphere: float = 1.0
    eps_sphere_center: List[int] = field(default_factory=lambda: [0, 0, 0])
    eps_sphere_radius: int = 0
    mu_sphere: float = 1.0
    mu_sphere_center: List[int] = field(default_factory=lambda: [0, 0, 0])
    mu_sphere_radius: int = 0
    omegape_sphere: float = 0.0
    omegape_sphere_center: List[int] = field(default_factory=lambda: [0, 0, 0])
    omegape_sphere_radius: int = 0
    omegapm_sphere: float = 0.0
    omegapm_sphere_center: List[int] = field(default_factory=lambda: [0, 0, 0])
    omegapm_sphere_radius: int = 0
    sphere_accuracy: int = 100
    stair_sphere: bool = False
    layout: int = 0

    @property
    def dim(self) -> int:
        return 3 if self.scheme == SCHEME_DIM3 else (2 if self.scheme >= SCHEME_TEX else 1)

    @property
    def exp_border(self) -> float:
        """The material border coordinate of the exp solutions: the eps sphere centre minus its
        radius along the pair's line axis (CallBack.cpp:297-760)."""
        axis = [v[1] for v in SCHEMES_1D.values() if v[0] == self.exp_pair][0]
        return float(self.eps_sphere_center[axis] - self.eps_sphere_radius)

    @property
    def active_axes(self):
        """The axes of the scheme's grid (its order: x before y before z)."""
        if self.dim == 1:
            return [v[1] for v in SCHEMES_1D.values() if v[0] == self.scheme]
        return [a for a in range(3) if a != self.missing_axis]

    @property
    def shape(self):
        if self.dim == 3:
            return tuple(self.size[:3])
        return tuple(self.size[a] for a in self.active_axes)

    @property
    def material_shape(self):
        """Shape of the material grids: the field grid's, doubled along every axis with
        --use-double-material-precision (YeeGridLayout sizeEps, YeeGridLayout.cpp:1972)."""
        return tuple(n * 2 for n in self.shape) if self.double_material else self.shape

    @property
    def missing_axis(self) -> int:
        """The axis a 2D scheme lacks (-1 in 3D and 1D)."""
        for scheme, axis in SCHEMES_2D.values():
            if scheme == self.scheme:
                return axis
        return -1

    @property
    def components(self):
        """Field components of the scheme (InternalScheme component flags): TE<a> has H along the
        missing axis a and E across it, TM<a> the reverse."""
        names = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]
        if self.dim == 3:
            return names
        if self.dim == 1:
            return [n for n in names if any(v[0] == self.scheme and n in v[2] for v in SCHEMES_1D.values())]
        m, te = self.missing_axis, self.scheme <= SCHEME_TEZ
        return [n for c, n in enumerate(names) if ((c >= 3) == te) == ((c % 3) == m)]

    def to_c(self) -> YeeConfigC:
        c = YeeConfigC()
        c.scheme = self.scheme
        c.layout = self.layout
        c.precision = self.precision
        c.time_steps = self.time_steps
        for a in range(3):
            c.size[a] = int(self.size[a])
            c.pml[a] = int(self.pml[a])
            c.tfsf_left[a] = int(self.tfsf_left[a])
            c.tfsf_right[a] = int(self.tfsf_right[a])
            c.angle_deg[a] = float(self.angle_deg[a])
            c.point_source_pos[a] = int(self.point_source_pos[a])
        c.current_source = 0
        for name in self.current_source:
            c.current_source |= 1 << COMP_INDEX[name]
        for a in range(3):
            c.current_source_pos[a] = int(self.current_source_pos[a])
        c.dx = self.dx
        c.wavelength = self.wavelength
        c.courant = self.courant
        c.flags = ((USE_PML if self.use_pml else 0) | (USE_TFSF if self.use_tfsf else 0)
                   | (USE_METAMATERIALS if self.use_metamaterials else 0)
                   | (COMPLEX_FIELDS if self.complex_fields else 0)
                   | (DOUBLE_MATERIAL_PRECISION if self.double_material else 0))
        mask = 0
        for name in self.point_source:
            mask |= 1 << COMP_INDEX[name]
        c.point_source = mask
        return c


class OptionError(ValueError):
    """Unknown or malformed option (the reference returns EXIT_UNKNOWN_OPTION, Settings.cpp:364-365)."""


def _parse_coord(text: str, default: int) -> List[int]:
    # Settings::parseCoordinate: comma-separated "x:<int>", "y:<int>", "z:<int>"
    vals = [default, default, default]
    for part in text.split(","):
        if ":" not in part:
            raise OptionError(f"bad coordinate '{text}'")
        axis, num = part.split(":", 1)
        idx = {"x": 0, "y": 1, "z": 2}.get(axis.strip())
        if idx is None:
            raise OptionError(f"bad coordinate axis in '{text}'")
        vals[idx] = int(num)
    return vals


# options that take one argument and are accepted but do not affect the Yee path
_IGNORED_WITH_ARG = {
    "--log-level", "--buffer-size", "--topology-size",
    "--cuda-buffer-size", "--cuda-gpus", "--num-cuda-threads", "--interm-save-step",
    "--save-start-coord", "--save-end-coord", "--rebalance-step", "--stored-steps",
    "--save-cmd-to-file",
}
_IGNORED_FLAGS = {
    "--same-size-topology", "--manual-topology", "--use-cuda",
    "--save-res", "--save-res-per-process", "--save-materials", "--save-interm-res",
    "--save-scattered-field-res", "--save-scattered--field-interm", "--save-as-bmp", "--save-as-dat",
    "--save-as-txt", "--save-tfsf-e-incident", "--save-tfsf-h-incident", "--manual-save-coords",
    "--save-no-pml", "--palette-gray", "--palette-rgb", "--orth-axis-x", "--orth-axis-y",
    "--orth-axis-z", "--dynamic-grid", "--check-disabling-cond", "--check-enabling-cond",
    # precomputed coefficient grids (Scheme.inc.h:3309-3791): the same expressions as the
    # per-cell path the engine computes, so the results are the same with or without them
    "--use-ca-cb", "--use-ca-cb-pml",
}
# options of the reference whose feature is outside this engine: rejected loudly
_UNSUPPORTED = {
    # the reference's precomputed Drude path reads a1 / a2 grids its binding never sets
    # (InternalScheme.inc.h:1902-2295, :2501-2502; Scheme.inc.h:932)
    "--use-ca-cb-pml-metamaterials",
}


# --use-<kind>-start-values / -border-condition / -right-side, --calc-<kind>-diff-norm
# (Settings.inc.h:134-153; sin1 has no right side)
_EXP_FN = re.compile(r"^--calc-(exp1|exp2|exp3)-(exhy|exhz|eyhx|eyhz|ezhx|ezhy)-diff-norm$")
_TEST_FN = re.compile(r"^--(use|calc)-(polinom1|polinom2|polinom3|sin1)-"
                      r"(start-values|border-condition|right-side|diff-norm)$")


def _read_cmd_file(path: str) -> List[str]:
    # Settings::loadCmdFromFile: one token per line, '#' and '//' comment lines
    out = []
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line or line.startswith("#") or line.startswith("//"):
                continue
            out.append(line)
    return out


def parse_args(argv: Sequence[str], precision: int = 8) -> Config:
    """Parse the reference's command line (subset that drives the Yee path)."""
    cfg = Config(precision=precision)
    args = list(argv)
    i = 0
    while i < len(args):
        a = args[i]

        def val():
            nonlocal i
            i += 1
            if i >= len(args):
                raise OptionError(f"option {a} needs an argument")
            return args[i]

        if a == "--cmd-from-file":
            rest = _read_cmd_file(val())
            args = args[: i + 1] + rest + args[i + 1:]
        elif a == "--3d":
            cfg.scheme = SCHEME_DIM3
        elif a == "--current-source":
            cfg.current_source_pos = _parse_coord(val(), 0)
        elif a.startswith("--current-source-") and a[-2:] in ("jx", "jy", "jz", "mx", "my", "mz"):
            cfg.current_source.append(("E" if a[-2] == "j" else "H") + a[-1])
        elif a in SCHEMES_2D:
            cfg.scheme = SCHEMES_2D[a][0]
        elif a in SCHEMES_1D:
            cfg.scheme = SCHEMES_1D[a][0]
        elif a == "--size":
            cfg.size = _parse_coord(val(), 100)
        elif a == "--same-size":
            cfg.size = [cfg.size[0]] * 3
        elif a == "--pml-size":
            cfg.pml = _parse_coord(val(), 2)
        elif a == "--same-size-pml":
            cfg.pml = [cfg.pml[0]] * 3
        elif a == "--tfsf-size-left":
            cfg.tfsf_left = _parse_coord(val(), 4)
        elif a == "--tfsf-size-right":
            cfg.tfsf_right = _parse_coord(val(), 4)
        elif a == "--same-size-tfsf":
            cfg.tfsf_left = [cfg.tfsf_left[0]] * 3
            cfg.tfsf_right = [cfg.tfsf_right[0]] * 3
        elif a == "--time-steps":
            cfg.time_steps = int(val())
        elif a == "--angle-teta":
            cfg.angle_deg[0] = stof(val())
        elif a == "--angle-phi":
            cfg.angle_deg[1] = stof(val())
        elif a == "--angle-psi":
            cfg.angle_deg[2] = stof(val())
        elif a == "--dx":
            cfg.dx = stof(val())
        elif a == "--wavelength":
            cfg.wavelength = stof(val())
        elif a == "--courant-factor":
            cfg.courant = stof(val())
        elif a == "--use-tfsf":
            cfg.use_tfsf = True
        elif a == "--use-pml":
            cfg.use_pml = True
        elif a == "--use-metamaterials":
            cfg.use_metamaterials = True
        elif a == "--layout-type":
            cfg.layout = int(val())
        elif a == "--point-source":
            cfg.point_source_pos = _parse_coord(val(), 0)
        elif a.startswith("--point-source-") and a[len("--point-source-"):] in ("ex", "ey", "ez", "hx", "hy", "hz"):
            comp = a[len("--point-source-"):]
            cfg.point_source.append(comp[0].upper() + comp[1])
        elif a == "--load-eps-from-file":
            cfg.eps_file = val()
        elif a == "--load-mu-from-file":
            cfg.mu_file = val()
        elif a == "--load-omegape-from-file":
            cfg.omegape_file = val()
        elif a == "--load-omegapm-from-file":
            cfg.omegapm_file = val()
        elif a == "--load-gammae-from-file":
            cfg.gammae_file = val()
        elif a == "--load-gammam-from-file":
            cfg.gammam_file = val()
        elif a == "--eps-sphere":
            cfg.eps_sphere = stof(val())
        elif a == "--eps-sphere-center":
            cfg.eps_sphere_center = _parse_coord(val(), 0)
        elif a == "--eps-sphere-radius":
            cfg.eps_sphere_radius = int(val())
        elif a == "--mu-sphere":
            cfg.mu_sphere = stof(val())
        elif a == "--mu-sphere-center":
            cfg.mu_sphere_center = _parse_coord(val(), 0)
        elif a == "--mu-sphere-radius":
            cfg.mu_sphere_radius = int(val())
        elif a == "--omegape-sphere":
            cfg.omegape_sphere = stof(val())
        elif a == "--omegape-sphere-center":
            cfg.omegape_sphere_center = _parse_coord(val(), 0)
        elif a == "--omegape-sphere-radius":
            cfg.omegape_sphere_radius = int(val())
        elif a == "--omegapm-sphere":
            cfg.omegapm_sphere = stof(val())
        elif a == "--omegapm-sphere-center":
            cfg.omegapm_sphere_center = _parse_coord(val(), 0)
        elif a == "--omegapm-sphere-radius":
            cfg.omegapm_sphere_radius = int(val())
        elif a == "--sphere-accuracy":
            cfg.sphere_accuracy = int(val())
        elif a == "--stair-sphere-approx":
            cfg.stair_sphere = True
        elif a == "--use-double-material-precision":
            cfg.double_material = True
        elif _TEST_FN.match(a) and ((_TEST_FN.match(a).group(1) == "calc") == (_TEST_FN.match(a).group(3) == "diff-norm")) \
                and a != "--use-sin1-right-side":
            m = _TEST_FN.match(a)
            kind = TEST_FUNCTIONS[m.group(2)]
            slot = {"start-values": "start_fn", "border-condition": "border_fn", "right-side": "rhs_fn",
                    "diff-norm": "exact_fn"}[m.group(3)]
            # Scheme::initCallBacks takes the first kind of polinom1, polinom2, polinom3, sin1
            cur = getattr(cfg, slot)
            setattr(cfg, slot, kind if cur == 0 else min(cur, kind))
        elif a == "--norm-start":
            cfg.norm_start = _parse_coord(val(), 0)
        elif a == "--norm-end":
            cfg.norm_end = _parse_coord(val(), 0)
        elif _EXP_FN.match(a):
            m = _EXP_FN.match(a)
            kind, pair = EXP_FUNCTIONS[m.group(1)], SCHEMES_1D["--1d-" + m.group(2)][0]
            if cfg.exp_fn and cfg.exp_pair != pair:
                raise OptionError("exact solutions of two different 1D pairs are not supported")
            # Scheme::initCallBacks takes exp1 before exp2 before exp3 within a pair
            cfg.exp_fn = kind if cfg.exp_fn == 0 else min(cfg.exp_fn, kind)
            cfg.exp_pair = pair
        elif a == "--use-ntff":
            cfg.use_ntff = True
        elif a == "--ntff-size":
            cfg.ntff_size = _parse_coord(val(), 3)
        elif a == "--same-size-ntff":
            cfg.ntff_size = [cfg.ntff_size[0]] * 3
        elif a == "--ntff-diff":
            cfg.ntff_diff = int(val())
        elif a == "--ntff-step-angle":
            cfg.ntff_step_angle = stof(val())
        elif a == "--ntff-filename":
            cfg.ntff_filename = val()
        elif a == "--ntff-to-stdout":
            cfg.ntff_to_stdout = True
        elif a == "--ntff-reverse":
            cfg.ntff_reverse = True
        elif a == "--interm-ntff-step":
            cfg.interm_ntff_step = int(val())
        elif a == "--eps-normed":
            cfg.eps_normed = True
        elif a == "--mu-normed":
            cfg.mu_normed = True
        elif a in _IGNORED_WITH_ARG:
            val()
        elif a in _IGNORED_FLAGS:
            pass
        elif a in _UNSUPPORTED:
            raise OptionError(f"option {a} is outside the Yee engine (3D, TMz and TEz, E_CENTERED, real fields)")
        else:
            raise OptionError(f"unknown option {a}")
        i += 1
    if cfg.use_metamaterials and not cfg.use_pml:
        raise OptionError("--use-metamaterials requires --use-pml (reference: Scheme.inc.h:889-892)")
    if cfg.use_metamaterials and cfg.double_material:
        raise OptionError("--use-double-material-precision excludes --use-metamaterials (the reference's "
                          "8-cell Drude average is UNREACHABLE, Approximation.cpp:196)")
    if cfg.layout not in (0, 1):
        raise OptionError("--layout-type must be 0 (E_CENTERED) or 1 (H_CENTERED)")
    if cfg.layout == 1 and (cfg.use_metamaterials or cfg.double_material):
        raise OptionError("the H_CENTERED layout (--layout-type 1) runs without --use-metamaterials and "
                          "--use-double-material-precision")
    return cfg
