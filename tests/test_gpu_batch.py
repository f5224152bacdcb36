"""Batched TMz sweep (yee_batch) against single runs and the oracle (needs the GPU).

Every member of a batch runs the same cell code as a single TMz handle, so the bar is
bit-identical fields against the single-run engine (FP32 and FP64), and the FP64 batch
against the oracle at the fixture bar (<= 1e-12 relative L2; the engine updates E/H
directly outside the PML slabs).
"""
import sys

import numpy as np
import pytest

from fdtd3d_amd.config import parse_args
from fdtd3d_amd.engine import Batch, Engine
from oracle.binding import Oracle
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def blobs(shape, seed):
    rng = np.random.default_rng(seed)
    x, y = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    eps = np.ones(shape)
    for _ in range(4):
        cx, cy = rng.uniform(0, shape[0]), rng.uniform(0, shape[1])
        r = rng.uniform(2, 6)
        eps += 8 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * r * r))
    return np.clip(eps, 1, 9)


SIMS = [
    ("--2d-tmz --size x:70,y:50 --use-pml --pml-size x:6,y:6 --point-source x:30,y:21 --point-source-ez", 40.0, 12.0),
    ("--2d-tmz --size x:33,y:129 --use-pml --pml-size x:5,y:7 --point-source x:16,y:60 --point-source-ez "
     "--point-source-hy", 25.0, 8.0),
    ("--2d-tmz --size x:100,y:64 --point-source x:50,y:1 --point-source-hx", 30.0, 10.0),
    ("--2d-tmz --size x:20,y:20 --use-pml --pml-size x:3,y:3 --point-source x:10,y:10 --point-source-ez", 0.0, 0.0),
]


def make(precision, steps):
    cfgs = []
    for args, _, _ in SIMS:
        c = parse_args((args + f" --time-steps {steps}").split())
        c.precision = precision
        cfgs.append(c)
    return cfgs


@pytest.mark.parametrize("precision", [4, 8])
@pytest.mark.parametrize("steps", [9, 61])
def test_batch_members_match_single_runs(precision, steps, engine_lib):
    cfgs = make(precision, steps)
    say(f"[gpu] batch vs single precision {precision} steps {steps}")
    b = Batch(cfgs)
    for s, (c, (_, t0, sig)) in enumerate(zip(cfgs, SIMS)):
        b.set_material(s, "eps", blobs(c.shape, 100 + s))
        b.set_point_source_gaussian(s, t0, sig)
    b.prepare()
    b.step(steps)
    b.sync()
    assert b.time == steps
    for s, (c, (_, t0, sig)) in enumerate(zip(cfgs, SIMS)):
        e = Engine(c)
        e.set_material("eps", blobs(c.shape, 100 + s))
        e.set_point_source_gaussian(t0, sig)
        e.prepare()
        e.step(steps)
        e.sync()
        for f in ("Ez", "Hx", "Hy"):
            assert np.array_equal(b.field(s, f), e.field(f)), f"sim {s} {f}"
        e.close()
    if steps >= 61:
        tun = b.tuning()
        assert tun["tile_rows"] in (13, 24, 40, 72) and tun["workgroups"] > 0 and tun["groups"] >= 1
        assert all(v > 0 for v in tun["candidate_ms"])
    b.close()


def test_batch_fp64_vs_oracle(engine_lib):
    steps = 80
    cfgs = make(8, steps)
    say("[gpu] batch vs oracle")
    b = Batch(cfgs)
    for s, (c, (_, t0, sig)) in enumerate(zip(cfgs, SIMS)):
        b.set_material(s, "eps", blobs(c.shape, 100 + s))
        b.set_point_source_gaussian(s, t0, sig)
    b.prepare()
    b.step(steps)
    for s, (c, (_, t0, sig)) in enumerate(zip(cfgs, SIMS)):
        o = Oracle(c)
        o.set_material("eps", blobs(c.shape, 100 + s))
        o.set_source_gaussian(t0, sig)
        o.prepare()
        o.step(steps)
        for f in ("Ez", "Hx", "Hy"):
            r = rel_l2(b.field(s, f), o.field(f))
            assert r <= TOL_FP64, f"sim {s} {f}: {r}"
    b.close()


def test_batch_steps_in_several_calls(engine_lib):
    """Graph replays, single steps and the tuner split over calls give the same state."""
    steps = [1, 30, 5, 17, 40]
    cfgs = make(4, sum(steps))
    say("[gpu] batch steps in several calls")
    a = Batch(cfgs)
    b = Batch(cfgs)
    for s, (c, (_, t0, sig)) in enumerate(zip(cfgs, SIMS)):
        for x in (a, b):
            x.set_material(s, "eps", blobs(c.shape, 100 + s))
            x.set_point_source_gaussian(s, t0, sig)
    a.prepare()
    b.prepare()
    for n in steps:
        a.step(n)
    b.step(sum(steps))
    for s in range(len(cfgs)):
        for f in ("Ez", "Hx", "Hy"):
            assert np.array_equal(a.field(s, f), b.field(s, f)), f"sim {s} {f}"
    a.close()
    b.close()


# sims whose tiles split into several x chunks and column strips, with sources on the halo
# rows and columns of the two-step tiles (chunks of 13 rows start at x = 1, 14, 27, ...;
# strips of 61 columns at y = 1, 62, 123, ...)
SIMS_EDGE = [
    ("--2d-tmz --size x:60,y:150 --use-pml --pml-size x:6,y:8 --point-source x:14,y:61 --point-source-ez", 20.0, 9.0),
    ("--2d-tmz --size x:45,y:130 --use-pml --pml-size x:5,y:5 --point-source x:13,y:62 --point-source-ez "
     "--point-source-hx --point-source-hy", 15.0, 6.0),
    ("--2d-tmz --size x:31,y:130 --point-source x:27,y:122 --point-source-ez", 10.0, 5.0),
    ("--2d-tmz --size x:29,y:140 --use-pml --pml-size x:4,y:4 --point-source x:26,y:123 --point-source-hy", 12.0, 4.0),
]

# the two-step launch shapes of the tuner (index into its candidate table: grouping,
# two steps, tile length) and one single-step shape
CANDS_TWO = [3, 4, 6, 9]


@pytest.mark.parametrize("precision", [4, 8])
@pytest.mark.parametrize("cand", CANDS_TWO + [0])
def test_batch_launch_shapes_match_single_runs(precision, cand, engine_lib, monkeypatch):
    steps = 24   # one source step, a graph of 16 steps, three two-step launches, one single step
    cfgs = []
    for args, _, _ in SIMS_EDGE:
        c = parse_args((args + f" --time-steps {steps}").split())
        c.precision = precision
        cfgs.append(c)
    say(f"[gpu] batch launch shape {cand} precision {precision}")
    monkeypatch.setenv("YEE_BATCH_CAND", str(cand))
    b = Batch(cfgs)
    for s, (c, (_, t0, sig)) in enumerate(zip(cfgs, SIMS_EDGE)):
        b.set_material(s, "eps", blobs(c.shape, 200 + s))
        b.set_point_source_gaussian(s, t0, sig)
    b.prepare()
    b.step(steps)
    b.sync()
    monkeypatch.delenv("YEE_BATCH_CAND")
    for s, (c, (_, t0, sig)) in enumerate(zip(cfgs, SIMS_EDGE)):
        e = Engine(c)
        e.set_material("eps", blobs(c.shape, 200 + s))
        e.set_point_source_gaussian(t0, sig)
        e.prepare()
        e.step(steps)
        e.sync()
        for f in ("Ez", "Hx", "Hy"):
            assert np.array_equal(b.field(s, f), e.field(f)), f"sim {s} {f}"
        if c.use_pml:
            for f in ("Dz", "Bx", "By"):
                assert np.array_equal(b.field(s, f), e.field(f)), f"sim {s} {f}"
        e.close()
    b.close()
