"""The oracle at the full 80^3 FP64 configuration (accurate eps = 4 sphere of radius 20 at
(40, 40, 40), PML 10, TF/SF inset 20, theta = phi = psi = 0, 200 steps) against the summary
the reference's own arithmetic wrote for it (tests/golden/cfg1_sphere80.npz: per component L2,
sum, largest magnitude, the mid-z plane and 4096 seeded cells; the eps grid).  Bit for bit.
About 80 s (the oracle is a plain C loop).
"""
import numpy as np

from oracle.binding import Oracle, sphere_map
from tests.helpers import load_summary

COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def test_oracle_cfg1_full_size_bit_exact_against_reference_summary():
    cfg, data = load_summary("cfg1_sphere80")
    assert cfg.size == [80, 80, 80] and cfg.time_steps == 200 and cfg.pml == [10, 10, 10]
    shape = cfg.shape
    eps = sphere_map(cfg, cfg.eps_sphere, 1.0, cfg.eps_sphere_center, cfg.eps_sphere_radius)
    np.testing.assert_array_equal(eps.ravel(), data["Eps"])   # partial-volume sphere, 100^2 samples per cell
    o = Oracle(cfg)
    o.set_material("eps", eps)
    o.prepare()
    o.step(cfg.time_steps)
    np.testing.assert_array_equal(o.scalars()[:2], data["scalars"][:2])
    for c in COMPS:
        v = o.field(c)
        assert np.array_equal(v[:, :, shape[2] // 2], data[c + "_midz"]), c
        assert np.array_equal(v.ravel()[data["cells"]], data[c + "_cells"]), c
        assert float(np.sqrt(np.sum(v * v))) == float(data[c + "_l2"]), c
        assert float(np.sum(v)) == float(data[c + "_sum"]), c
        assert float(np.abs(v).max()) == float(data[c + "_max"]), c
