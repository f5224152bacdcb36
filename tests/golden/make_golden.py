"""Writes the golden fixtures of tests/golden/ from the reference's own arithmetic.

Runs oracle/_ref/refdrv_f64 (built by `make -C oracle ref` from the reference sources;
only possible where the reference tree is mounted) on small inputs and stores the
newest time layer of every field, the material and sigma grids, the head of the
incident 1D grids and the run scalars, as compressed .npz files.

    python tests/golden/make_golden.py            # regenerate all fixtures
    python tests/golden/make_golden.py cfg1_sphere80   # one fixture

cfg1_sphere80 is the 80^3 FP64 configuration in full (accurate eps = 4 sphere of radius 20,
PML 10, TF/SF inset 20, 200 steps).  Its fields are too large to commit whole, so it stores a
summary: per component the L2 norm, sum and largest magnitude, the full mid-z plane (z = 40), and
4096 cells drawn with a fixed seed; plus the eps grid (mostly 1 and 4: it compresses to a few
kB).  It runs the reference driver's -O3 build (no asserts; the same arithmetic, built with the
same -ffp-contract=off), about two minutes.
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.binding import run_ref_driver  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

W0 = 2 * np.pi * 299792458.0 / np.float64(np.float32(0.02))


def drude_files(tmp, n, lo, hi):
    om = np.zeros((n, n, n))
    ga = np.zeros((n, n, n))
    om[lo:hi, lo:hi, lo:hi] = 1.5 * W0
    ga[lo:hi, lo:hi, lo:hi] = 0.01 * W0
    p_om = os.path.join(tmp, "omega.dat")
    p_ga = os.path.join(tmp, "gamma.dat")
    om.tofile(p_om)
    ga.tofile(p_ga)
    return p_om, p_ga


def cases(tmp):
    """name -> (argv, material files written for the run)"""
    out = {}
    out["d3_pml_tfsf_oblique_sphere"] = (
        "--3d --size x:16 --same-size --dx 0.0005 --wavelength 0.02 --courant-factor 0.5 "
        "--use-pml --pml-size x:3 --same-size-pml --use-tfsf --tfsf-size-left x:4 --tfsf-size-right x:4 "
        "--same-size-tfsf --angle-teta 30 --angle-phi 45 --angle-psi 90 --eps-sphere 4 "
        "--eps-sphere-center x:8,y:8,z:8 --eps-sphere-radius 4 --time-steps 30").split()
    out["d3_tfsf_axial_pointsource"] = (
        "--3d --size x:16,y:14,z:18 --use-tfsf --tfsf-size-left x:4,y:3,z:5 --tfsf-size-right x:3,y:4,z:4 "
        "--angle-teta 0 --angle-phi 0 --angle-psi 0 --point-source x:3,y:4,z:5 --point-source-hy "
        "--time-steps 25").split()
    om, ga = drude_files(tmp, 16, 6, 10)
    out["d3_drude_pml_tfsf"] = (
        "--3d --size x:16 --same-size --use-pml --pml-size x:3 --same-size-pml --use-tfsf "
        "--tfsf-size-left x:4 --tfsf-size-right x:4 --same-size-tfsf --angle-teta 0 --angle-phi 0 "
        f"--angle-psi 0 --use-metamaterials --load-omegape-from-file {om} --load-omegapm-from-file {om} "
        f"--load-gammae-from-file {ga} --load-gammam-from-file {ga} --time-steps 30").split()
    out["tmz_pml_pointsource_sphere"] = (
        "--2d-tmz --size x:32,y:24 --use-pml --pml-size x:4,y:5 --point-source x:13,y:11 --point-source-ez "
        "--eps-sphere 3 --eps-sphere-center x:16,y:12 --eps-sphere-radius 5 --time-steps 50").split()
    out["tmz_pml_tfsf_oblique"] = (
        "--2d-tmz --size x:32,y:28 --use-pml --pml-size x:4,y:4 --use-tfsf --tfsf-size-left x:7,y:8 "
        "--tfsf-size-right x:6,y:7 --angle-teta 90 --angle-phi 30 --angle-psi 90 --time-steps 50").split()
    # TEz: the TEz layout asserts theta = 90 and psi = 0 (YeeGridLayout.cpp:2442)
    out["tez_pml_tfsf_oblique_sphere"] = (
        "--2d-tez --size x:32,y:28 --use-pml --pml-size x:4,y:4 --use-tfsf --tfsf-size-left x:7,y:8 "
        "--tfsf-size-right x:6,y:7 --angle-teta 90 --angle-phi 30 --angle-psi 0 --eps-sphere 3 "
        "--eps-sphere-center x:16,y:14 --eps-sphere-radius 5 --time-steps 50").split()
    out["tez_pml_pointsource"] = (
        "--2d-tez --size x:30,y:26 --use-pml --pml-size x:4,y:5 --point-source x:13,y:11 --point-source-hz "
        "--point-source-ex --angle-psi 0 --time-steps 50").split()
    # the (y, z) and (x, z) planes: TMx / TEx assert phi = 90, TMy / TEy phi = 0
    out["tmx_pml_tfsf_sphere"] = (
        "--2d-tmx --size x:7,y:30,z:26 --use-pml --pml-size y:4,z:4 --use-tfsf --tfsf-size-left y:7,z:6 "
        "--tfsf-size-right y:6,z:8 --angle-teta 35 --angle-phi 90 --angle-psi 0 --eps-sphere 3 "
        "--eps-sphere-center x:3,y:15,z:13 --eps-sphere-radius 5 --time-steps 50").split()
    out["tex_pml_tfsf_pointsource"] = (
        "--2d-tex --size x:5,y:28,z:24 --use-pml --pml-size y:3,z:4 --use-tfsf --tfsf-size-left y:7,z:6 "
        "--tfsf-size-right y:6,z:7 --angle-teta 55 --angle-phi 90 --angle-psi 90 --point-source x:0,y:12,z:10 "
        "--point-source-hx --point-source-ez --time-steps 50").split()
    out["tmy_pml_tfsf_pointsource"] = (
        "--2d-tmy --size x:27,y:7,z:25 --use-pml --pml-size x:4,z:3 --use-tfsf --tfsf-size-left x:7,z:6 "
        "--tfsf-size-right x:6,z:7 --angle-teta 40 --angle-phi 0 --angle-psi 0 --point-source x:12,y:0,z:11 "
        "--point-source-ey --time-steps 50").split()
    out["tey_pml_tfsf_sphere"] = (
        "--2d-tey --size x:25,y:6,z:29 --use-pml --pml-size x:3,z:4 --use-tfsf --tfsf-size-left x:6,z:7 "
        "--tfsf-size-right x:7,z:6 --angle-teta 25 --angle-phi 0 --angle-psi 90 --eps-sphere 2.5 "
        "--eps-sphere-center x:12,y:2,z:14 --eps-sphere-radius 5 --time-steps 50").split()
    # 1D lines (main.cpp:629-800): ExHy along z with a Drude layer inside PML + TF/SF; EzHy along
    # x with a point source and J / M currents
    n1 = 120
    om1, ga1 = np.zeros(n1), np.zeros(n1)
    om1[70:85], ga1[70:85] = 1.5 * W0, 0.01 * W0
    p_om1, p_ga1 = os.path.join(tmp, "omega1d.dat"), os.path.join(tmp, "gamma1d.dat")
    om1.tofile(p_om1)
    ga1.tofile(p_ga1)
    out["exhy_pml_tfsf_drude"] = (
        f"--1d-exhy --size z:{n1} --use-pml --pml-size z:10 --use-tfsf --tfsf-size-left z:20 "
        f"--tfsf-size-right z:20 --angle-teta 0 --angle-phi 0 --angle-psi 90 --use-metamaterials "
        f"--load-omegape-from-file {p_om1} --load-omegapm-from-file {p_om1} --load-gammae-from-file {p_ga1} "
        f"--load-gammam-from-file {p_ga1} --time-steps 300").split()
    out["ezhy_pml_sources"] = (
        "--1d-ezhy --size x:90 --use-pml --pml-size x:8 --point-source x:30 --point-source-ez "
        "--current-source x:55 --current-source-jz --current-source-my --angle-teta 90 --angle-phi 0 "
        "--angle-psi 90 --time-steps 200").split()
    # the H_CENTERED layout (--layout-type 1; Tests/suite t1.2, t2.3, t6.x, t7.x and t8 run both
    # layouts): 3D with and without PML, the TMz / TEz / TEy planes and two 1D lines
    hc = ["--layout-type", "1"]
    out["hc_d3_pml_tfsf_oblique_sphere"] = out["d3_pml_tfsf_oblique_sphere"] + hc
    out["hc_d3_tfsf_axial_pointsource"] = out["d3_tfsf_axial_pointsource"] + hc
    out["hc_tmz_pml_tfsf_oblique"] = out["tmz_pml_tfsf_oblique"] + hc
    out["hc_tez_pml_tfsf_oblique_sphere"] = out["tez_pml_tfsf_oblique_sphere"] + hc
    out["hc_tey_pml_tfsf_sphere"] = out["tey_pml_tfsf_sphere"] + hc
    out["hc_exhy_pml_tfsf"] = (
        f"--1d-exhy --size z:{n1} --use-pml --pml-size z:10 --use-tfsf --tfsf-size-left z:20 "
        "--tfsf-size-right z:20 --angle-teta 0 --angle-phi 0 --angle-psi 90 --eps-sphere 2 "
        "--eps-sphere-center z:75 --eps-sphere-radius 10 --time-steps 300").split() + hc
    out["hc_ezhy_pml_sources"] = out["ezhy_pml_sources"] + hc
    return out


def complex_cases():
    """name -> argv of the fixtures written by the complex-field build (refdrv_c64): the real parts
    under the component names, the imaginary parts as "<name>.im"."""
    return {
        "c3d_pml_tfsf_sources": (
            "--3d --size x:16,y:15,z:17 --use-pml --pml-size x:3,y:3,z:3 --use-tfsf --tfsf-size-left x:5,y:5,z:5 "
            "--tfsf-size-right x:5,y:5,z:5 --angle-teta 30 --angle-phi 45 --angle-psi 90 --eps-sphere 3 "
            "--eps-sphere-center x:8,y:7,z:8 --eps-sphere-radius 3 --point-source x:7,y:7,z:9 --point-source-hy "
            "--current-source x:9,y:8,z:8 --current-source-jx --current-source-mz --time-steps 30").split(),
        "ctmz_pml_tfsf_current": (
            "--2d-tmz --size x:30,y:26 --use-pml --pml-size x:4,y:4 --use-tfsf --tfsf-size-left x:7,y:7 "
            "--tfsf-size-right x:7,y:7 --angle-teta 90 --angle-phi 30 --angle-psi 90 --current-source x:14,y:12 "
            "--current-source-jz --current-source-mx --point-source x:10,y:9 --point-source-ez --time-steps 50").split(),
        # the H_CENTERED layout with complex sources
        "hc_c3d_pml_tfsf_sources": (
            "--3d --size x:16,y:15,z:17 --use-pml --pml-size x:3,y:3,z:3 --use-tfsf --tfsf-size-left x:5,y:5,z:5 "
            "--tfsf-size-right x:5,y:5,z:5 --angle-teta 30 --angle-phi 45 --angle-psi 90 --eps-sphere 3 "
            "--eps-sphere-center x:8,y:7,z:8 --eps-sphere-radius 3 --point-source x:7,y:7,z:9 --point-source-hy "
            "--current-source x:9,y:8,z:8 --current-source-jx --current-source-mz --time-steps 30 "
            "--layout-type 1").split(),
    }


# Text dumps written by the reference's own TXTDumper (the reference driver with YEE_REF_TXT),
# FP64 without PML, so the engine's fields are bit-identical to the reference's and its CLI's
# --save-as-txt dumps must match these files byte for byte.
TXT_DUMP_ARGV = ("--3d --size x:10,y:9,z:11 --use-tfsf --tfsf-size-left x:3 --tfsf-size-right x:3 --same-size-tfsf "
                 "--angle-teta 30 --angle-phi 45 --angle-psi 90 --eps-sphere 3 --eps-sphere-center x:5,y:4,z:5 "
                 "--eps-sphere-radius 3 --point-source x:4,y:4,z:6 --point-source-hz --time-steps 20")
TXT_DUMP_COMPONENTS = ("Ex", "Hz")


def write_txt_dumps():
    out_dir = os.path.join(HERE, "txt_dumps_3d")
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        os.environ["YEE_REF_TXT"] = tmp
        try:
            run_ref_driver(TXT_DUMP_ARGV.split(), os.path.join(tmp, "t.bin"))
        finally:
            del os.environ["YEE_REF_TXT"]
        for c in TXT_DUMP_COMPONENTS:
            with open(os.path.join(tmp, c + ".txt"), "rb") as src, open(os.path.join(out_dir, c + ".txt"), "wb") as dst:
                dst.write(src.read())
    with open(os.path.join(out_dir, "argv.txt"), "w") as fh:
        fh.write(TXT_DUMP_ARGV + "\n")
    print("txt_dumps_3d", TXT_DUMP_COMPONENTS)


CFG1_ARGV = ("--3d --size x:80 --same-size --dx 0.0005 --wavelength 0.02 --angle-teta 0 --angle-phi 0 "
             "--angle-psi 0 --use-tfsf --tfsf-size-left x:20 --tfsf-size-right x:20 --same-size-tfsf --use-pml "
             "--pml-size x:10 --same-size-pml --courant-factor 0.5 --time-steps 200 --eps-sphere 4 "
             "--eps-sphere-center x:40,y:40,z:40 --eps-sphere-radius 20")
SUMMARY_SEED = 0xC0F1
SUMMARY_CELLS = 4096


def summary_payload(rec: dict, shape, argv: str) -> dict:
    """Per component: L2, sum, max |v|, the mid-z plane and SUMMARY_CELLS seeded cells."""
    rng = np.random.default_rng(SUMMARY_SEED)
    flat = rng.choice(int(np.prod(shape)), size=SUMMARY_CELLS, replace=False)
    out = {"argv": np.array(argv), "cells": flat.astype(np.int64), "Eps": rec["Eps"],
           "scalars": rec["scalars"]}
    for name in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        v = rec[name]
        out[name + "_l2"] = np.array(np.sqrt(np.sum(v * v)))
        out[name + "_sum"] = np.array(np.sum(v))
        out[name + "_max"] = np.array(np.abs(v).max())
        out[name + "_midz"] = v.reshape(shape)[:, :, shape[2] // 2].copy()
        out[name + "_cells"] = v[flat]
    return out


def write_cfg1(bin_path: str = ""):
    """cfg1_sphere80 from the reference driver (or from an existing driver output file)."""
    from oracle.binding import read_ref_output
    with tempfile.TemporaryDirectory() as tmp:
        if bin_path:
            rec = read_ref_output(bin_path)
        else:
            rec, _ = run_ref_driver(CFG1_ARGV.split(), os.path.join(tmp, "cfg1.bin"), kind="f64_fast", timeout=3600)
    payload = summary_payload(rec, (80, 80, 80), CFG1_ARGV)
    np.savez_compressed(os.path.join(HERE, "cfg1_sphere80.npz"), **payload)
    print("cfg1_sphere80", sorted(payload))


def main():
    only = set(a for a in sys.argv[1:] if not a.startswith("--"))   # optional: the fixture names to (re)generate
    if not only or "txt_dumps_3d" in only:
        write_txt_dumps()
        only.discard("txt_dumps_3d")
        if not only and len(sys.argv) > 1:
            return
    if not only or "cfg1_sphere80" in only:
        write_cfg1(next((a.split("=", 1)[1] for a in sys.argv[1:] if a.startswith("--from=")), ""))
        only.discard("cfg1_sphere80")
        if not only and len(sys.argv) > 1:
            return
    with tempfile.TemporaryDirectory() as tmp:
        for name, argv in complex_cases().items():
            if only and name not in only:
                continue
            rec, _ = run_ref_driver(argv, os.path.join(tmp, name + ".bin"), kind="c64")
            payload = {"argv": np.array(" ".join(argv)), "complex": np.array(True)}
            for key, val in rec.items():
                if key.startswith("EInc") or key.startswith("HInc"):
                    val = val[:256]
                payload[key] = val
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **payload)
            print(name, sorted(payload))
        for name, argv in cases(tmp).items():
            if only and name not in only:
                continue
            rec, _ = run_ref_driver(argv, os.path.join(tmp, name + ".bin"))
            payload = {"argv": np.array(" ".join(a if not a.startswith(tmp) else os.path.basename(a) for a in argv))}
            for key, val in rec.items():
                if key in ("EInc", "HInc"):
                    val = val[:256]
                payload[key] = val
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **payload)
            print(name, sorted(payload))


if __name__ == "__main__":
    main()
