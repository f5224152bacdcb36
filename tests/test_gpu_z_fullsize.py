"""Parity at the benchmark sizes (needs the GPU), through properties that do not depend on an
oracle run, which would take hours at these sizes:

  - 1024^3 FP32 (the headline workload) and 768^3 FP64: the fused step kernel and the
    half-step kernels compute the same expressions in the same order, so after a few steps
    every field is bit-identical between the two paths (chunking, XCD ordering, z-face
    staging and both workgroup shapes at their real sizes);
  - 1024^3: FP32 within the 1e-5 relative-L2 bar of FP64 on the same (FP32-rounded) inputs,
    the FP64 engine being pinned to the oracle at 1e-12 on smaller grids;
  - the 64-simulation TMz batch: each simulation bit-identical to the same simulation run
    alone;
  - the 256^3 adjoint: dJ/deps against 4th-order central differences of the FP64 forward
    run at the 1e-4 bar.

Each test holds one or two handles of up to ~200 GB of device memory at a time and copies
fields to host memory component by component.
"""
import sys

import numpy as np
import pytest

import bench
from fdtd3d_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu

COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run_cfg2(cfg, kernel, steps, monkeypatch):
    monkeypatch.setenv("YEE_KERNEL", kernel)
    eng = Engine(cfg)
    eng.set_material("eps", bench.block_eps(cfg.shape, dtype=np.float32))
    eng.prepare()
    eng.step(steps)
    eng.sync()
    return eng


AUX = ("Dx", "Dy", "Dz", "Bx", "By", "Bz")


@pytest.mark.parametrize("n,precision", [(1024, 4), (768, 8)])
def test_fullsize_fused_matches_half_steps(n, precision, engine_lib, monkeypatch):
    cfg = bench.cfg2(n=n, precision=precision)
    # the incident wave enters the total-field box after about 5 steps; the field scattered by
    # the eps blocks crosses the 8 cells between the TF/SF box and the PML slabs in about 20
    steps = 48
    say(f"[gpu] full size {n}^3 precision {precision}: half-step kernels, {steps} steps")
    eng = _run_cfg2(cfg, "split", steps, monkeypatch)
    ref = {c: eng.field_raw(c) for c in COMPS + AUX}
    eng.close()
    say(f"[gpu] full size {n}^3 precision {precision}: fused kernel, {steps} steps")
    eng = _run_cfg2(cfg, "fused", steps, monkeypatch)
    for c in COMPS + AUX:
        got = eng.field_raw(c)
        assert np.abs(ref[c]).max() > 0, c   # the wave has reached every component (D / B: the PML slabs)
        same = np.array_equal(got, ref[c])
        if not same:
            diff = np.argwhere(got != ref[c])
            pytest.fail(f"{c}: {len(diff)} cells differ, first at {tuple(diff[0])}")
        ref[c] = None
        del got
    eng.close()


def test_fullsize_fp32_within_bar_of_fp64(engine_lib):
    class A:
        workload = "cfg2"
        check_steps = 100
    say("[gpu] full size 1024^3 FP32 vs FP64, 100 steps")
    out = bench.precision_check(A)
    say(f"[gpu] rel L2 per component: {out['rel_l2']}")
    for c, r in out["rel_l2"].items():
        assert r <= 1e-5, (c, r)


def test_fullsize_batch_matches_single_runs(engine_lib):
    steps = 60
    sims = bench.cfg4_batch(steps=steps)
    say(f"[gpu] full size batch: {len(sims)} TMz simulations, {steps} steps")
    b = Batch([c for c, _ in sims])
    for s, (c, blobs) in enumerate(sims):
        b.set_material(s, "eps", bench.blob_eps(c.shape, blobs))
        b.set_point_source_gaussian(s, 10.0, 5.0)
    b.prepare()
    b.step(steps)
    b.sync()
    for s, (c, blobs) in enumerate(sims):
        e = Engine(c)
        e.set_material("eps", bench.blob_eps(c.shape, blobs))
        e.set_point_source_gaussian(10.0, 5.0)
        e.prepare()
        e.step(steps)
        e.sync()
        for name in ("Ez", "Hx", "Hy"):
            got = b.field(s, name)
            want = e.field(name)
            assert np.abs(want).max() > 0, (s, name)
            assert np.array_equal(got, want), (s, name, float(np.abs(got - want).max()))
        e.close()
    b.close()


def test_fullsize_adjoint_gradient_vs_central_differences(engine_lib):
    c, spec, design = bench.cfg5_adjoint(precision=8)
    say(f"[gpu] full size adjoint {c.size} FP64, {c.time_steps} steps")
    eps = np.ones(c.shape)
    lo, hi = spec["design_lo"], spec["design_hi"]
    eps[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = design
    e = Engine(c)
    e.set_material("eps", eps)
    e.prepare()
    e.adjoint_setup(spec["design_lo"], spec["design_hi"], spec["monitor_z"], spec["monitor_lo"], spec["monitor_hi"],
                    spec["window"], checkpoint_every=50)
    e.forward(design)
    grad = e.backward(1.0)
    rng = np.random.default_rng(0x5EED)
    for _ in range(3):
        v = tuple(int(rng.integers(0, design.shape[a])) for a in range(3))
        h = 1e-2 * design[v]
        js = []
        for m in (2, 1, -1, -2):
            d = design.copy()
            d[v] += m * h
            js.append(e.forward(d))
        fd = (-js[0] + 8 * js[1] - 8 * js[2] + js[3]) / (12 * h)
        rel = abs(grad[v] - fd) / abs(fd)
        say(f"[gpu] voxel {v}: adjoint {grad[v]:.9e} fd {fd:.9e} rel {rel:.2e}")
        assert rel <= 1e-4, (v, grad[v], fd)
    e.close()
