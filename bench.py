#!/usr/bin/env python3
"""Benchmark of the Yee time-stepping engine on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg2]

Default workload (the configuration the headline metric is quoted on): 3D 1024^3 FP32,
PML 16 cells, TF/SF plane wave lambda = 40 cells at theta = 30, phi = 45 (psi = 90),
Courant 0.5, eps piecewise constant over 8^3 blocks drawn U[1, 12] from a fixed seed,
mu = 1.  A "step" is one Yee time step of the whole grid; the timed region runs K steps
with every input resident in HBM.  Prints one JSON line.

roofline: algorithmic bytes per cell-update B_alg = w (12 + 12 f_pml + 36 f_drude) + 1
(every state word read and written once per step, plus a 1-byte material class), times
the cells of one step, divided by the device time of one step measured with HIP events
on the engine's stream over the timed region, against the 8 TB/s HBM3E peak.

cpu_baseline: the reference's own per-cell arithmetic (oracle/_ref/refdrv_f32_fast, built
from the reference sources) timed on one host core on a bounded sample of the same
workload; the C oracle stands in where that binary is absent.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fdtd3d_amd.config import Config, SCHEME_DIM3, SCHEME_TMZ, stof  # noqa: E402

METRIC = "3D Yee Mcell-updates/s + % HBM roofline (1 GPU); adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
EPS_SEED = 0xFD7D3D


def cfg2(n=1024, precision=4, pml=16, inset=24, steps=1000) -> Config:
    return Config(scheme=SCHEME_DIM3, precision=precision, time_steps=steps, size=[n, n, n], pml=[pml] * 3,
                  tfsf_left=[inset] * 3, tfsf_right=[inset] * 3, dx=stof("0.0005"), wavelength=stof("0.02"),
                  courant=stof("0.5"), angle_deg=[stof("30"), stof("45"), stof("90")], use_pml=True, use_tfsf=True)


def block_eps(shape, block=8, lo=1.0, hi=12.0, seed=EPS_SEED, dtype=np.float32):
    """eps piecewise constant over block^3 cubes, i.i.d. U[lo, hi] per cube."""
    rng = np.random.default_rng(seed)
    nb = [(s + block - 1) // block for s in shape]
    vals = rng.uniform(lo, hi, size=nb).astype(dtype)
    out = np.repeat(np.repeat(np.repeat(vals, block, axis=0), block, axis=1), block, axis=2)
    return np.ascontiguousarray(out[: shape[0], : shape[1], : shape[2]])


WORKLOADS = {
    "cfg2": dict(desc="3D 1024^3 FP32, PML 16, TF/SF lambda=40 cells theta=30 phi=45 psi=90, Courant 0.5, "
                      "eps 8^3 blocks U[1,12] seed 0xFD7D3D, mu=1",
                 make=lambda: cfg2()),
    "cfg2-768-fp64": dict(desc="3D 768^3 FP64, otherwise as cfg2", make=lambda: cfg2(n=768, precision=8)),
    "cfg2-512": dict(desc="3D 512^3 FP32, otherwise as cfg2", make=lambda: cfg2(n=512)),
    "cfg2-512-hc": dict(desc="3D 512^3 FP32, H_CENTERED layout (--layout-type 1, half-step kernels), otherwise as cfg2",
                        make=lambda: _hc(cfg2(n=512))),
    "cfg2-plain": dict(desc="3D 1024^3 FP32 without PML and TF/SF (stencil only), eps as cfg2",
                       make=lambda: _plain(cfg2())),
    "cfg3-drude": dict(desc="3D 512^3 FP32, PML 16, TF/SF lambda=40 cells theta=phi=psi=0, central 128^3 Drude slab "
                            "omegaPE=omegaPM=1.5 w0, gammaE=gammaM=0.01 w0 (w0 = 2 pi c / lambda), eps=mu=1",
                       make=lambda: cfg3()),
}


def cfg3(n=512, steps=2000) -> Config:
    return Config(scheme=SCHEME_DIM3, precision=4, time_steps=steps, size=[n, n, n], pml=[16] * 3,
                  tfsf_left=[24] * 3, tfsf_right=[24] * 3, dx=stof("0.0005"), wavelength=stof("0.02"),
                  courant=stof("0.5"), angle_deg=[0.0, 0.0, 0.0], use_pml=True, use_tfsf=True,
                  use_metamaterials=True)


def drude_materials(cfg: Config) -> dict:
    """Central 128^3 Drude slab (rad/s), zero elsewhere."""
    n = cfg.size[0]
    w0 = 2 * math.pi * 299792458.0 / cfg.wavelength
    lo, hi = n // 2 - 64, n // 2 + 64
    om = np.zeros(cfg.shape, dtype=np.float32)
    ga = np.zeros(cfg.shape, dtype=np.float32)
    om[lo:hi, lo:hi, lo:hi] = 1.5 * w0
    ga[lo:hi, lo:hi, lo:hi] = 0.01 * w0
    return {"omegape": om, "omegapm": om, "gammae": ga, "gammam": ga}


def cfg4_batch(nsim=64, seed=0xB47C4, steps=2000, precision=4):
    """64 TMz simulations: edges log-uniform in [256, 2048], PML 12, Gaussian-modulated hard Ez
    source at a random interior cell, eps = 1 + 8 * sum of Gaussian blobs clamped to [1, 9]."""
    rng = np.random.default_rng(seed)
    sims = []
    for s in range(nsim):
        nx, ny = (int(round(math.exp(rng.uniform(math.log(256), math.log(2048))))) for _ in range(2))
        px, py = int(rng.integers(12 + 1, nx - 12 - 1)), int(rng.integers(12 + 1, ny - 12 - 1))
        c = Config(scheme=SCHEME_TMZ, precision=precision, time_steps=steps, size=[nx, ny, 1], pml=[12, 12, 0],
                   use_pml=True, point_source=["Ez"], point_source_pos=[px, py, 0])
        blobs = [(rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(8, 40)) for _ in range(int(rng.integers(4, 10)))]
        sims.append((c, blobs))
    return sims


def blob_eps(shape, blobs, dtype=np.float32):
    x = np.arange(shape[0], dtype=np.float64)[:, None]
    y = np.arange(shape[1], dtype=np.float64)[None, :]
    eps = np.ones(shape)
    for cx, cy, r in blobs:
        eps += 8 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * r * r))
    return np.clip(eps, 1, 9).astype(dtype)


# source envelope of the batched sweep: one period of the source is
# lambda / (c dt) = (cells per wavelength) / Courant = 40 / 0.5 = 80 steps
GAUSS_SIGMA = 80.0
GAUSS_T0 = 4 * GAUSS_SIGMA


def run_cfg4(args) -> dict:
    """Batched sweep vs the same simulations run one by one (both on the GPU, same kernels)."""
    from fdtd3d_amd.engine import Batch, Engine
    sims = cfg4_batch(steps=args.batch_steps)
    cfgs = [c for c, _ in sims]
    cells = sum(c.size[0] * c.size[1] for c in cfgs)
    upd = cells * (args.batch_steps - 1)
    print(f"[bench] cfg4: {len(cfgs)} TMz simulations, {cells} cells, {args.batch_steps} steps", file=sys.stderr, flush=True)
    b = Batch(cfgs)
    for s, (c, blobs) in enumerate(sims):
        b.set_material(s, "eps", blob_eps(c.shape, blobs))
        b.set_point_source_gaussian(s, GAUSS_T0, GAUSS_SIGMA)
    b.prepare()
    b.sync()
    t0 = time.perf_counter()
    b.step(args.batch_steps)
    b.sync()
    t_batch = time.perf_counter() - t0
    tun = b.tuning()
    print(f"[bench] batched: {t_batch:.3f} s ({b.last_step_ms():.4f} ms/step); tuning {tun}", file=sys.stderr, flush=True)
    b.close()
    t_seq = 0.0
    if args.batch_only:
        return {"cells": cells, "updates": upd, "t_batch": t_batch, "t_seq": None, "tuning": tun,
                "ms_per_step": t_batch * 1e3 / args.batch_steps}
    for s, (c, blobs) in enumerate(sims):
        e = Engine(c)
        e.set_material("eps", blob_eps(c.shape, blobs))
        e.set_point_source_gaussian(GAUSS_T0, GAUSS_SIGMA)
        e.prepare()
        e.sync()
        t0 = time.perf_counter()
        e.step(args.batch_steps)
        e.sync()
        t_seq += time.perf_counter() - t0
        e.close()
    print(f"[bench] sequential: {t_seq:.3f} s", file=sys.stderr, flush=True)
    return {"cells": cells, "updates": upd, "t_batch": t_batch, "t_seq": t_seq, "tuning": tun,
            "ms_per_step": t_batch * 1e3 / args.batch_steps}


def cfg5_adjoint(precision=4, n=256, steps=800):
    """256^3, PML 16, TF/SF plane wave lambda = 32 cells (dx 0.000625, lambda 0.02), 800 steps;
    design 64^3 centred, eps U[1, 4] per voxel (seeded); J = sum of Ez^2 over the 128 x 128
    centre z-plane over the last 200 steps."""
    c = Config(scheme=SCHEME_DIM3, precision=precision, time_steps=steps, size=[n, n, n], pml=[16] * 3,
               tfsf_left=[24] * 3, tfsf_right=[24] * 3, dx=stof("0.000625"), wavelength=stof("0.02"),
               courant=stof("0.5"), use_pml=True, use_tfsf=True)
    lo, hi = n // 2 - 32, n // 2 + 32
    spec = dict(design_lo=(lo, lo, lo), design_hi=(hi, hi, hi), monitor_z=n // 2, monitor_lo=(n // 2 - 64, n // 2 - 64),
                monitor_hi=(n // 2 + 64, n // 2 + 64), window=200)
    design = np.random.default_rng(0xAD701).uniform(1.0, 4.0, size=(hi - lo,) * 3)
    return c, spec, design


def run_cfg5(args) -> dict:
    """Forward + backward time over plain forward time; FP64 gradient vs central differences."""
    from fdtd3d_amd.engine import Engine
    out = {}
    c, spec, design = cfg5_adjoint(precision=4)
    # plain forward: the same run without recording
    e = Engine(c)
    eps = np.ones(c.shape, dtype=np.float32)
    lo, hi = spec["design_lo"], spec["design_hi"]
    eps[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = design
    e.set_material("eps", eps)
    e.prepare()
    e.step(c.time_steps)
    e.sync()
    out["plain_forward_ms"] = e.last_step_ms() * c.time_steps
    e.close()
    for ck in (50, 0):
        e = Engine(c)
        e.set_material("eps", eps)
        e.prepare()
        e.adjoint_setup(spec["design_lo"], spec["design_hi"], spec["monitor_z"], spec["monitor_lo"],
                        spec["monitor_hi"], spec["window"], checkpoint_every=ck)
        J = e.forward(design)
        g32 = e.backward(1.0)
        tm = e.adjoint_timing()
        key = f"checkpoint{ck}" if ck else "trace"
        out[key] = {"J": J, "forward_ms": round(tm["forward_ms"], 2), "backward_ms": round(tm["backward_ms"], 2),
                    "ratio": round((tm["forward_ms"] + tm["backward_ms"]) / out["plain_forward_ms"], 3)}
        print(f"[bench] cfg5 {key}: {out[key]}", file=sys.stderr, flush=True)
        e.close()
    # FP64: gradient vs central finite differences on 16 seeded voxels, and the FP32 gradient error
    c64, _, _ = cfg5_adjoint(precision=8)
    e = Engine(c64)
    e.set_material("eps", eps.astype(np.float64))
    e.prepare()
    e.adjoint_setup(spec["design_lo"], spec["design_hi"], spec["monitor_z"], spec["monitor_lo"], spec["monitor_hi"],
                    spec["window"], checkpoint_every=50)
    e.forward(design)
    g64 = e.backward(1.0)
    rng = np.random.default_rng(0xFD16)
    errs, errs2 = [], []

    def central(v, h):
        dp, dm = design.copy(), design.copy()
        dp[v] += h
        dm[v] -= h
        return (e.forward(dp) - e.forward(dm)) / (2 * h)

    for _ in range(args.fd_voxels):
        v = tuple(int(rng.integers(0, design.shape[a])) for a in range(3))
        # h = 1e-2 eps: J(+h) - J(-h) must stand well above the rounding of J (about 1e-13 of
        # J after 800 FP64 steps); the h^2 error of the plain central difference is then
        # removed by the Richardson combination of the differences at h and 2h, the 4th-order
        # central difference (-J(+2h) + 8 J(+h) - 8 J(-h) + J(-2h)) / 12h
        h = 1e-2 * design[v]
        d1, d2 = central(v, h), central(v, 2 * h)
        fd = (4 * d1 - d2) / 3
        errs.append(abs(g64[v] - fd) / max(abs(fd), 1e-30))
        errs2.append(abs(g64[v] - d1) / max(abs(d1), 1e-30))
        print(f"[bench] fd voxel {v}: adjoint {g64[v]:.6e} fd4 {fd:.6e} rel {errs[-1]:.2e} (2nd order {errs2[-1]:.2e})",
              file=sys.stderr, flush=True)
    e.close()
    out["fd_max_rel_err_fp64"] = float(max(errs)) if errs else None
    out["fd2_max_rel_err_fp64"] = float(max(errs2)) if errs2 else None
    out["fd_voxels"] = len(errs)
    out["fp32_vs_fp64_grad_rel_l2"] = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    return out


def adjoint_ratio(fd_check: bool = False) -> dict:
    """(forward + backward) / plain forward of the 256^3 FP32 adjoint workload, checkpoint every
    50 steps, from HIP events on the engine stream."""
    from fdtd3d_amd.engine import Engine
    c, spec, design = cfg5_adjoint(precision=4)
    eps = np.ones(c.shape, dtype=np.float32)
    lo, hi = spec["design_lo"], spec["design_hi"]
    eps[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = design
    e = Engine(c)
    e.set_material("eps", eps)
    e.prepare()
    e.step(c.time_steps)
    e.sync()
    plain_ms = e.last_step_ms() * c.time_steps
    e.close()
    e = Engine(c)
    e.set_material("eps", eps)
    e.prepare()
    e.adjoint_setup(spec["design_lo"], spec["design_hi"], spec["monitor_z"], spec["monitor_lo"], spec["monitor_hi"],
                    spec["window"], checkpoint_every=50)
    e.forward(design)
    e.backward(1.0)
    tm = e.adjoint_timing()
    e.close()
    return {"value": round((tm["forward_ms"] + tm["backward_ms"]) / plain_ms, 3), "target": "<= 3",
            "plain_forward_ms": round(plain_ms, 1), "forward_ms": round(tm["forward_ms"], 1),
            "backward_ms": round(tm["backward_ms"], 1),
            "workload": "256^3 FP32, PML 16, TF/SF lambda 32 cells, 800 steps, design 64^3, checkpoint every 50 "
                        "steps (gradient checks: tests/test_gpu_z_fullsize.py)"}


def default_extras(args) -> dict:
    """The metric's second figure (adjoint/forward time ratio) and the batch speed-up, on their
    own workloads, appended to the default line."""
    out = {}
    try:
        t0 = time.perf_counter()
        out["adjoint_ratio"] = adjoint_ratio()
        print(f"[bench] adjoint ratio {out['adjoint_ratio']} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
              flush=True)
    except Exception as exc:
        out["adjoint_ratio"] = {"error": str(exc)[:200]}
    try:
        t0 = time.perf_counter()

        class A:
            batch_steps = 2000
            batch_only = False
        r = run_cfg4(A)
        out["batch_speedup"] = {"value": round(r["t_seq"] / r["t_batch"], 2), "target": ">= 3",
                                "batched_s": round(r["t_batch"], 3), "sequential_s": round(r["t_seq"], 3),
                                "mcell_updates_per_s": round(r["updates"] / r["t_batch"] / 1e6, 1),
                                "workload": "64 TMz sims, edges log-uniform [256, 2048], PML 12, 2000 steps, FP32"}
        print(f"[bench] batch speed-up {out['batch_speedup']['value']} ({time.perf_counter() - t0:.1f} s)",
              file=sys.stderr, flush=True)
    except Exception as exc:
        out["batch_speedup"] = {"error": str(exc)[:200]}
    return out


def precision_check(args) -> dict:
    """FP32 against FP64 on the same inputs (the FP32-rounded materials), both on the GPU, after
    --check-steps time steps: relative L2 of every field component.  The FP64 engine is itself
    pinned to the oracle at 1e-12, so this is the FP32 error of the full-size run."""
    from fdtd3d_amd.engine import Engine
    wl = WORKLOADS[args.workload]
    steps = args.check_steps
    ref = {}
    out = {}
    for prec in (8, 4):
        cfg = wl["make"]()
        cfg.precision = prec
        cfg.time_steps = steps
        eng = Engine(cfg)
        if args.workload == "cfg3-drude":
            for name, arr in drude_materials(cfg).items():
                eng.set_material(name, arr)
        else:
            eng.set_material("eps", block_eps(cfg.shape, dtype=np.float32))
        eng.prepare()
        t0 = time.perf_counter()
        eng.step(steps)
        eng.sync()
        print(f"[bench] precision check: FP{8 * prec} {steps} steps in {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        comps = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
        for comp in comps:
            f = eng.field_raw(comp)
            if prec == 8:
                ref[comp] = f
                continue
            num = den = 0.0
            for x0 in range(0, f.shape[0], 64):   # in slabs of planes: no full-size temporaries
                r = ref[comp][x0:x0 + 64]
                d = f[x0:x0 + 64].astype(np.float64) - r
                num += float(np.vdot(d, d))
                den += float(np.vdot(r, r))
            out[comp] = math.sqrt(num / den) if den > 0 else 0.0
            del f
            ref[comp] = None
        eng.close()
    return {"steps": steps, "rel_l2": out, "max_rel_l2": max(out.values())}


def _plain(c: Config) -> Config:
    c.use_pml = False
    c.use_tfsf = False
    return c


def _hc(c: Config) -> Config:
    c.layout = 1
    return c


def f_pml(cfg: Config) -> float:
    inner = 1.0
    for a in range(cfg.dim):
        inner *= max(0.0, 1.0 - 2.0 * cfg.pml[a] / cfg.size[a])
    return 1.0 - inner if cfg.use_pml else 0.0


def b_alg(cfg: Config, f_drude: float = 0.0) -> float:
    w = cfg.precision
    if cfg.dim == 3:
        return w * (12 + 12 * f_pml(cfg) + 36 * f_drude) + 1
    return w * (6 + 6 * f_pml(cfg)) + 1


def fused_cells(cfg: Config) -> int:
    """Cells the fused step kernel updates: the PML-free box minus the TF/SF x/y face cells."""
    n = cfg.size
    lo = [max(1, cfg.pml[a] if cfg.use_pml else 0) for a in range(3)]
    hi = [min(n[a] - 1, n[a] - (cfg.pml[a] if cfg.use_pml else 0) - 1) for a in range(3)]
    ext = [max(0, hi[a] - lo[a]) for a in range(3)]
    total = ext[0] * ext[1] * ext[2]
    if not cfg.use_tfsf:
        return total
    def span(a):   # [L-1, R] within the box
        L, R = cfg.tfsf_left[a], n[a] - cfg.tfsf_right[a]
        return max(0, min(R + 1, hi[a]) - max(L - 1, lo[a]))
    def faces(a):
        L, R = cfg.tfsf_left[a], n[a] - cfg.tfsf_right[a]
        return sum(1 for p in {L - 1, R - 1, R} if lo[a] <= p < hi[a])
    near = faces(0) * span(1) * span(2) + span(0) * faces(1) * span(2) - faces(0) * faces(1) * span(2)
    return total - near


def pmc_traffic(workload: str):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 PMC summary."""
    if workload != "cfg2":
        return None
    path = os.path.join(ROOT, "profiles", "step_3d_fused_cfg2_1024_fp32_hbm_traffic.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def host_cpu() -> dict:
    """CPU model and logical core count of the host the baseline runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def cpu_baseline(sample_n: int = 96, sample_steps: int = 31) -> dict:
    """Reference arithmetic on one core, config-2 features at sample_n^3 (T-1 updates)."""
    from oracle.binding import ref_driver_path, run_ref_driver
    cells = sample_n ** 3
    upd = cells * (sample_steps - 1)
    sample = (f"{sample_n}^3 FP32, PML 16, TF/SF inset 24, theta=30 phi=45 psi=90, lambda=40 cells, "
              f"8^3-block eps U[1,12], {sample_steps} time steps ({sample_steps - 1} field updates)")
    drv = ref_driver_path("f32_fast")
    if os.path.exists(drv):
        with tempfile.TemporaryDirectory() as tmp:
            epsf = os.path.join(tmp, "eps.dat")
            block_eps((sample_n,) * 3, dtype=np.float32).tofile(epsf)
            argv = (f"--3d --size x:{sample_n} --same-size --dx 0.0005 --wavelength 0.02 --courant-factor 0.5 "
                    f"--use-pml --pml-size x:16 --same-size-pml --use-tfsf --tfsf-size-left x:24 "
                    f"--tfsf-size-right x:24 --same-size-tfsf --angle-teta 30 --angle-phi 45 --angle-psi 90 "
                    f"--load-eps-from-file {epsf} --time-steps {sample_steps}").split()
            _, secs = run_ref_driver(argv, os.path.join(tmp, "out.bin"), kind="f32_fast", timeout=600)
        return {"value": upd / secs / 1e6, "unit": "Mcell-updates/s", "cores": 1, "kind": "reference",
                "sample": sample + "; the reference's own InternalScheme per-cell code (default path, compiled from "
                          "its sources with g++ -O3) under the step loop of oracle/ref_driver.cpp, not the "
                          "reference CLI's 'Total time'", **host_cpu()}
    from fdtd3d_amd.config import Config as _C
    from oracle.binding import Oracle
    c = cfg2(n=sample_n, precision=8, steps=sample_steps)
    o = Oracle(c)
    o.set_material("eps", block_eps((sample_n,) * 3, dtype=np.float64))
    o.prepare()
    t0 = time.perf_counter()
    o.step(sample_steps)
    secs = time.perf_counter() - t0
    return {"value": upd / secs / 1e6, "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
            "sample": sample + "; C oracle FP64 (reference driver not built)", **host_cpu()}


def copy_probe(mib: int) -> None:
    """One device-to-device copy of `mib` MiB (torch is only the allocator here)."""
    import torch
    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.neg(src, out=dst)   # one elementwise kernel: reads mib MiB, writes mib MiB
    torch.cuda.synchronize()
    print(f"[bench] copy probe: {mib} MiB read + {mib} MiB written", file=sys.stderr, flush=True)
    del src, dst
    torch.cuda.empty_cache()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="cfg2", choices=sorted(WORKLOADS) + ["cfg4-batch", "cfg5-adjoint"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=2000, help="time steps of the cfg4-batch workload")
    ap.add_argument("--precision-check", action="store_true",
                    help="instead of timing: FP32 vs FP64 relative L2 per component after --check-steps steps")
    ap.add_argument("--check-steps", type=int, default=1000, help="time steps of --precision-check")
    ap.add_argument("--batch-only", action="store_true",
                    help="cfg4-batch: time the batched run only (for profiling; no speed-up figure)")
    ap.add_argument("--fd-voxels", type=int, default=16, help="finite-difference voxels of the cfg5-adjoint workload")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload: skip the adjoint/forward ratio and the batch speed-up measurements")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before the run, copy a buffer of this many MiB device-to-device (a plain copy "
                         "of known size that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the grid is not sharded across GPUs; see the sharding design notes)", file=sys.stderr)
        sys.exit(2)

    from fdtd3d_amd.engine import Engine, device_info

    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)

    if args.workload == "cfg4-batch":
        r = run_cfg4(args)
        line = {
            "metric": "batched TMz sweep: Mcell-updates/s and speed-up over the same simulations run one by one",
            "value": round(r["updates"] / r["t_batch"] / 1e6, 1), "unit": "Mcell-updates/s", "n_gpus": 1,
            "steps": args.batch_steps, "warmup": 0, "ms_per_step": round(r["ms_per_step"], 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "cfg4-batch", "simulations": 64, "cells": r["cells"],
                       "description": "64 TMz sims, edges log-uniform [256, 2048] seed 0xB47C4, PML 12, "
                                      "Gaussian-modulated Ez source (t0 = 320, sigma = 80 steps), eps blobs in [1, 9]"},
            "sequential_s": round(r["t_seq"], 3) if r["t_seq"] is not None else None,
            "batched_s": round(r["t_batch"], 3),
            "speedup_vs_sequential": round(r["t_seq"] / r["t_batch"], 2) if r["t_seq"] is not None else None,
            "tuning": r["tuning"],
        }
        print(json.dumps(line), flush=True)
        return

    if args.workload == "cfg5-adjoint":
        r = run_cfg5(args)
        line = {
            "metric": "adjoint/forward time ratio: (forward + backward) / plain forward, 3D 256^3 FP32",
            "value": r["checkpoint50"]["ratio"], "unit": "ratio", "n_gpus": 1, "steps": 800, "warmup": 0,
            "ms_per_step": round(r["plain_forward_ms"] / 800, 4), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "cfg5-adjoint", "description": "256^3 PML 16, TF/SF lambda 32 cells, 800 steps, "
                       "design 64^3 eps U[1,4] seed 0xAD701, J = sum Ez^2 on the 128x128 centre z-plane over "
                       "the last 200 steps, checkpoint every 50 steps"},
            "detail": r,
        }
        print(json.dumps(line), flush=True)
        return

    if args.precision_check:
        if args.workload not in WORKLOADS:
            sys.exit("--precision-check applies to the 3D workloads")
        r = precision_check(args)
        print(json.dumps({"workload": args.workload, "check": "fp32 vs fp64 on the GPU, same inputs", **r}), flush=True)
        return

    wl = WORKLOADS[args.workload]
    cfg = wl["make"]()
    cells = int(np.prod(cfg.shape))
    print(f"[bench] {device_info()}", file=sys.stderr, flush=True)
    print(f"[bench] workload {args.workload}: {wl['desc']}", file=sys.stderr, flush=True)
    t_setup = time.perf_counter()
    eng = Engine(cfg)
    if args.workload == "cfg3-drude":
        for name, arr in drude_materials(cfg).items():
            eng.set_material(name, arr)
        f_drude = (128 / cfg.size[0]) ** 3
    else:
        eps = block_eps(cfg.shape, dtype=np.float32 if cfg.precision == 4 else np.float64)
        eng.set_material("eps", eps)
        del eps
        f_drude = 0.0
    eng.prepare()
    print(f"[bench] set-up {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    eng.step(args.warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.sync()
    wall = time.perf_counter() - t0
    dev_ms = eng.last_step_ms()          # HIP events on the engine stream over the timed region
    ms_per_step = wall * 1e3 / args.steps
    value = cells * args.steps / wall / 1e6
    prof = eng.profile_steps(min(args.steps, 20))
    print(f"[bench] {args.steps} steps in {wall:.3f} s; device {dev_ms:.3f} ms/step; kernels {prof}",
          file=sys.stderr, flush=True)

    balg = b_alg(cfg, f_drude)
    step_achieved = balg * cells / (dev_ms * 1e-3) / 1e9
    kernel_ms = {k: round(v, 4) for k, v in prof.items()}
    if prof["fused_ms"] > 0:
        # dominant kernel: the fused E+H step over the PML-free box (no PML terms there:
        # B_alg = w * 12 + 1 per cell), its duration from HIP events around each launch
        fcells = fused_cells(cfg)
        fb = cfg.precision * 12 + 1
        achieved = fb * fcells / (prof["fused_ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(args.workload),
                    "kernel": "step_3d_fused", "cells_per_launch": fcells, "b_alg_bytes_per_cell_update": fb,
                    "kernel_ms": kernel_ms}
    else:
        roofline = {"bound": "hbm", "achieved": round(step_achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(step_achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "one time step (e_half_3d + h_half_3d)", "kernel_ms": kernel_ms}
    # the whole step against the same roofline (B_alg of the metric: PML slabs included)
    roofline["step"] = {"achieved": round(step_achieved, 1), "frac": round(step_achieved / HBM_PEAK_GBS, 4),
                        "b_alg_bytes_per_cell_update": round(balg, 2), "ms": round(dev_ms, 4)}
    eng.close()

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline()
        except Exception as exc:  # the GPU figure stands on its own
            cpu = {"error": str(exc)[:200]}

    extras = {}
    if args.workload == "cfg2" and not args.no_extras:
        extras = default_extras(args)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if cfg.precision == 4 else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "description": wl["desc"], "grid": list(cfg.shape),
                   "pml": cfg.pml[0], "tfsf_inset": cfg.tfsf_left[0], "time_steps_of_run": cfg.time_steps,
                   "cells": cells},
        "roofline": roofline, "cpu_baseline": cpu, **extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
