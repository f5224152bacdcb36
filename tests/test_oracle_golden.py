"""The C oracle against golden fixtures written from the reference's own arithmetic.

The fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py with
oracle/_ref/refdrv_f64) hold the newest time layer of every field after a short run
of 3D and TMz cases with PML, TF/SF at oblique incidence, Drude metamaterials, point
sources and partial-volume spheres.  The oracle must reproduce them bit for bit.
"""
import numpy as np
import pytest

from oracle.binding import Oracle, run_oracle_complex, sphere_map
from tests.helpers import field_names, golden_names, load_golden

# The H_CENTERED fixtures (hc_*) pin the engine directly: the oracle restates the E_CENTERED
# layout only.  Their own check here: the layout changed the fields.
NAMES = [n for n in golden_names() if not n.startswith("hc_")]
HC_NAMES = [n for n in golden_names() if n.startswith("hc_")]


@pytest.mark.parametrize("name", HC_NAMES)
def test_h_centered_fixture_differs_from_e_centered(name):
    cfg, _, data = load_golden(name)
    assert cfg.layout == 1
    base = name[3:]
    if base not in NAMES:
        pytest.skip("no E_CENTERED twin")
    _, _, ref = load_golden(base)
    assert any(not np.array_equal(data[f], ref[f]) for f in field_names(cfg))


def test_fixtures_present():
    assert len(NAMES) >= 5


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference_fixture_bit_exact(name):
    cfg, mats, data = load_golden(name)
    if cfg.complex_fields:
        # the complex-field build: its real and imaginary parts are the oracle's runs with the
        # complex sources' real and imaginary parts
        parts = run_oracle_complex(cfg, mats)
        for fname in field_names(cfg, aux=True):
            for o, key in zip(parts, (fname, fname + ".im")):
                got = o.field(fname).ravel()
                assert np.array_equal(got, data[key]), f"{name}: {key} differs (max {np.abs(got - data[key]).max()})"
        for o, sfx in zip(parts, ("", ".im")):
            e, h = o.incident()
            np.testing.assert_array_equal(e[:256], data["EInc" + sfx])
            np.testing.assert_array_equal(h[:256], data["HInc" + sfx])
        return
    o = Oracle(cfg)
    for key, arr in mats.items():
        o.set_material(key, arr)
    o.prepare()
    o.step(cfg.time_steps)
    for fname in field_names(cfg, aux=True):
        got = o.field(fname).ravel()
        assert np.array_equal(got, data[fname]), f"{name}: {fname} differs (max {np.abs(got - data[fname]).max()})"
    np.testing.assert_array_equal(o.scalars()[:2], data["scalars"][:2])
    if cfg.use_tfsf:
        e, h = o.incident()
        np.testing.assert_array_equal(e[:256], data["EInc"])
        np.testing.assert_array_equal(h[:256], data["HInc"])
    if cfg.use_pml:
        for axis, key in enumerate(["SigmaX", "SigmaY", "SigmaZ"]):
            if axis in cfg.active_axes:   # no sigma along an axis the scheme lacks
                np.testing.assert_array_equal(o.material(6 + axis).ravel(), data[key])


@pytest.mark.parametrize("name", NAMES)
def test_sphere_generator_matches_reference(name):
    cfg, mats, data = load_golden(name)
    if cfg.eps_sphere == 1.0:
        pytest.skip("no sphere in this fixture")
    eps = sphere_map(cfg, cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius)
    np.testing.assert_array_equal(eps.ravel(), data["Eps"])
