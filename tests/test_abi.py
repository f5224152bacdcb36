"""C-ABI checks that need no GPU: the library loads and exports every declared entry point,
and the option parser follows the reference's option table."""
import os
import re

import pytest

from fdtd3d_amd.config import OptionError, parse_args, stof, SCHEME_TMZ, SCHEME_DIM3
from fdtd3d_amd.engine import LIB_PATH, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if hdr.endswith(".h"):
            text = open(os.path.join(ROOT, "include", hdr)).read()
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(yee_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_built():
    assert os.path.exists(LIB_PATH), "run python -m fdtd3d_amd.build"


def test_exports_every_declared_symbol():
    L = load_library()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_error_reporting_without_gpu():
    import ctypes
    from fdtd3d_amd.config import Config
    L = load_library()
    h = ctypes.c_void_p()
    bad = Config(scheme=14).to_c()
    assert L.yee_create(ctypes.byref(bad), ctypes.byref(h)) < 0
    assert b"scheme" in L.yee_last_error()
    assert L.yee_create(None, ctypes.byref(h)) < 0
    assert L.yee_step(None, 1) < 0


def test_parse_reference_options():
    cfg = parse_args("--3d --size x:80 --same-size --dx 0.0005 --wavelength 0.02 --angle-teta 30 "
                     "--angle-phi 45 --angle-psi 90 --use-tfsf --tfsf-size-left x:20 --tfsf-size-right x:20 "
                     "--same-size-tfsf --use-pml --pml-size x:10 --same-size-pml --courant-factor 0.5 "
                     "--time-steps 200 --eps-sphere 4 --eps-sphere-center x:40,y:40,z:40 --eps-sphere-radius 20".split())
    assert cfg.scheme == SCHEME_DIM3 and cfg.size == [80, 80, 80] and cfg.pml == [10, 10, 10]
    assert cfg.tfsf_left == [20, 20, 20] and cfg.tfsf_right == [20, 20, 20]
    assert cfg.dx == stof("0.0005") and cfg.dx != 0.0005   # std::stof rounding, as the reference
    assert cfg.use_pml and cfg.use_tfsf and cfg.time_steps == 200
    assert cfg.eps_sphere == 4 and cfg.eps_sphere_center == [40, 40, 40] and cfg.eps_sphere_radius == 20
    tmz = parse_args("--2d-tmz --size x:64,y:48 --point-source x:3,y:4 --point-source-ez".split())
    assert tmz.scheme == SCHEME_TMZ and tmz.shape == (64, 48) and tmz.point_source == ["Ez"]


def test_parse_rejects_unknown_and_unsupported():
    with pytest.raises(OptionError):
        parse_args(["--no-such-option"])
    with pytest.raises(OptionError):
        parse_args(["--3d", "--use-metamaterials"])
    with pytest.raises(OptionError):
        # the reference's precomputed Drude path dereferences a1 / a2 grids it never binds
        parse_args("--3d --use-pml --use-metamaterials --use-ca-cb-pml-metamaterials".split())


def test_parse_exp_solution_options():
    """--calc-exp<n>-<pair>-diff-norm (Settings.inc.h:153-175): kind and pair, exp1 before exp2
    before exp3 within a pair (Scheme::initCallBacks), two pairs refused; the dielectric border
    is the eps sphere centre minus its radius along the pair's line."""
    from fdtd3d_amd.config import EXP_FUNCTIONS, SCHEMES_1D
    c = parse_args("--1d-eyhz --size x:60 --eps-sphere 4 --eps-sphere-center x:40 --eps-sphere-radius 15 "
                   "--calc-exp3-eyhz-diff-norm --calc-exp2-eyhz-diff-norm".split())
    assert c.exp_fn == EXP_FUNCTIONS["exp2"] == 6 and c.exp_pair == SCHEMES_1D["--1d-eyhz"][0]
    assert c.exp_border == 25.0
    with pytest.raises(OptionError):
        parse_args("--1d-exhy --calc-exp1-exhy-diff-norm --calc-exp1-ezhy-diff-norm".split())


def test_parse_ntff_options():
    """--use-ntff and its settings (Settings.inc.h:85-88, 124, 181-187): defaults, --same-size-ntff,
    and the diagram's angles as Scheme::saveNTFF's loop adds them up."""
    import math
    from fdtd3d_amd.engine import ntff_angles
    d = parse_args(["--3d", "--use-ntff"])
    assert d.use_ntff and d.ntff_size == [3, 3, 3] and d.ntff_diff == 1 and d.ntff_step_angle == 10.0
    assert d.ntff_filename == "ntff-res" and d.interm_ntff_step == 100 and not d.ntff_reverse and not d.ntff_to_stdout
    c = parse_args("--3d --use-ntff --ntff-size x:5,y:4 --same-size-ntff --ntff-diff 3 --ntff-step-angle 7.5 "
                   "--ntff-filename far --ntff-to-stdout --ntff-reverse --interm-ntff-step 20".split())
    assert c.ntff_size == [5, 5, 5] and c.ntff_diff == 3 and c.ntff_step_angle == 7.5
    assert c.ntff_filename == "far" and c.ntff_to_stdout and c.ntff_reverse and c.interm_ntff_step == 20
    assert list(ntff_angles(c)) == [c.angle_deg[1] * math.pi / 180.0]
    fwd = ntff_angles(d)
    assert len(fwd) == 37 and fwd[0] == 0.0 and abs(fwd[-1] - 2 * math.pi) < 1e-12
    step = math.pi * 10.0 / 180
    acc = 0.0
    for a in fwd:
        assert a == acc
        acc += step
    # 1 degree: 361 additions of pi / 180 stay 4.6e-14 below 2 pi + pi / 180, so the loop also
    # prints that angle (362 lines)
    one = ntff_angles(parse_args("--3d --ntff-step-angle 1".split()))
    assert len(one) == 362 and one[-1] < 2 * math.pi + math.pi / 180


def test_parse_accepts_precomputed_coefficient_flags():
    """--use-ca-cb / --use-ca-cb-pml select the reference's precomputed coefficient grids, whose
    results equal the per-cell path (tests/test_oracle_ref_live.py): accepted, same config."""
    base = parse_args("--3d --size x:20 --same-size --use-pml --pml-size x:4 --same-size-pml".split())
    cacb = parse_args("--3d --size x:20 --same-size --use-pml --pml-size x:4 --same-size-pml --use-ca-cb "
                      "--use-ca-cb-pml".split())
    assert base == cacb


def test_cmd_file(tmp_path):
    f = tmp_path / "cmd.txt"
    f.write_text("# comment\n--3d\n--size\nx:12,y:10,z:8\n// another\n--time-steps\n7\n")
    cfg = parse_args(["--cmd-from-file", str(f)])
    assert cfg.size == [12, 10, 8] and cfg.time_steps == 7


def test_cli_rejects_options_outside_the_path():
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fdtd3d_amd", "bin", "fdtd3d")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    import subprocess
    for argv in (["--no-such-flag"], ["--size"], ["--use-ntff"],
                 ["--use-pml", "--use-metamaterials", "--use-ca-cb-pml-metamaterials"]):
        out = subprocess.run([exe] + argv, capture_output=True, text=True, timeout=60)
        assert out.returncode != 0, argv
        assert out.stderr.startswith("fdtd3d: "), out.stderr


INVALID_CONFIGS = [
    (dict(precision=2), b"precision"),
    (dict(layout=2), b"H_CENTERED"),
    (dict(size=[10, 10, 10], layout=1, use_pml=True, use_metamaterials=True), b"H_CENTERED"),
    (dict(size=[2, 10, 10]), b"size"),
    (dict(size=[10, 10, 1 << 21]), b"size"),
    (dict(size=[8, 8, 8], pml=[5, 1, 1], use_pml=True), b"pml"),
    (dict(size=[10, 10, 10], pml=[-1, 1, 1], use_pml=True), b"pml"),
    (dict(size=[10, 10, 10], use_tfsf=True, tfsf_left=[0, 3, 3], tfsf_right=[3, 3, 3]), b"TF/SF"),
    (dict(size=[10, 10, 10], use_tfsf=True, tfsf_left=[5, 3, 3], tfsf_right=[5, 3, 3]), b"TF/SF"),
    (dict(size=[4096, 4096, 2048]), b"too large"),
    (dict(size=[10, 10, 10], dx=0.0), b"positive"),
    (dict(size=[10, 10, 10], courant=-0.5), b"positive"),
    (dict(size=[10, 10, 10], use_metamaterials=True), b"PML"),
    (dict(size=[10, 10, 10], time_steps=-1), b"time steps"),
    (dict(size=[10, 10, 10], current_source=["Ez"], current_source_pos=[0, 5, 5]), b"current source"),
    (dict(scheme=1, size=[10, 10, 2]), b"size"),                      # ExHy: the line along z
    (dict(scheme=4, size=[20, 1, 1], pml=[11, 0, 0], use_pml=True), b"pml"),   # EyHz along x
    (dict(scheme=0, size=[10, 10, 10]), b"scheme"),
    (dict(size=[10, 10, 10], current_source=["Hx"], current_source_pos=[5, 9, 5]), b"current source"),
]


def test_parse_1d_schemes():
    from fdtd3d_amd.config import SCHEME_EXHY, SCHEME_EZHY
    c = parse_args("--1d-exhy --size z:80 --use-pml --pml-size z:6".split())
    assert c.scheme == SCHEME_EXHY and c.dim == 1 and c.shape == (80,) and c.components == ["Ex", "Hy"]
    c = parse_args("--1d-ezhy --size x:40".split())
    assert c.scheme == SCHEME_EZHY and c.active_axes == [0] and c.shape == (40,) and c.components == ["Ez", "Hy"]


def test_parse_current_sources():
    cfg = parse_args("--3d --size x:12 --same-size --current-source x:5,y:6,z:7 --current-source-jz "
                     "--current-source-mx".split())
    assert cfg.current_source == ["Ez", "Hx"] and cfg.current_source_pos == [5, 6, 7]
    c = cfg.to_c()
    assert c.current_source == (1 << 2) | (1 << 3) and list(c.current_source_pos) == [5, 6, 7]


@pytest.mark.parametrize("kw,msg", INVALID_CONFIGS)
def test_create_rejects_invalid_configs(kw, msg):
    """yee_create validates the configuration before touching the device: every invalid
    configuration returns YEE_EINVAL with a message naming the problem."""
    import ctypes
    from fdtd3d_amd.config import Config
    L = load_library()
    h = ctypes.c_void_p()
    c = Config(**kw).to_c()
    assert L.yee_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert msg in L.yee_last_error(), L.yee_last_error()
    assert not h.value


def test_parse_test_function_options():
    """--use-<kind>-* / --calc-<kind>-diff-norm take the first kind in the reference's priority
    (polinom1, polinom2, polinom3, sin1); sin1 has no right side; --norm-start / --norm-end and
    --eps-normed / --mu-normed are read."""
    from fdtd3d_amd.config import OptionError, parse_args
    cfg = parse_args("--3d --use-polinom3-border-condition --use-polinom2-border-condition --use-sin1-start-values "
                     "--use-polinom1-right-side --calc-sin1-diff-norm --norm-start x:2,z:3 --norm-end y:7 "
                     "--eps-normed --mu-normed".split())
    assert (cfg.border_fn, cfg.start_fn, cfg.rhs_fn, cfg.exact_fn) == (2, 4, 1, 4)
    assert cfg.norm_start == [2, 0, 3] and cfg.norm_end == [0, 7, 0]
    assert cfg.eps_normed and cfg.mu_normed
    for bad in ("--use-sin1-right-side", "--use-polinom1-diff-norm", "--calc-polinom1-border-condition"):
        with pytest.raises(OptionError):
            parse_args(["--3d", bad])
