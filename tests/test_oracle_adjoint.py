"""The FP64 CPU adjoint of the oracle against central finite differences of the forward
oracle (the adjoint has no reference counterpart; the finite differences pin it).

J = sum over the last `window` steps of sum of Ez^2 on a monitor rectangle; the gradient is
dJ/d eps over a design box.  Central differences with h = 1e-4 eps have a truncation error
of O(h^2) (about 1e-6 relative) and a rounding error of about 1e-16 J / h, below the 1e-5
bar used here.
"""
import numpy as np
import pytest

from fdtd3d_amd.config import parse_args
from oracle.binding import oracle_adjoint, oracle_objective

CASES = [
    # PML + oblique TF/SF plane wave, design box inside the total-field region
    ("--3d --size x:18,y:16,z:20 --use-pml --pml-size x:3,y:3,z:3 --use-tfsf --tfsf-size-left x:5,y:5,z:5 "
     "--tfsf-size-right x:5,y:5,z:5 --angle-teta 30 --angle-phi 45 --angle-psi 90 --wavelength 0.005 --time-steps 36",
     (7, 6, 8), (11, 10, 12), 13, (5, 5), (13, 11), 12),
    # no PML, point source, design box touching the monitor plane
    ("--3d --size x:14,y:14,z:14 --point-source x:4,y:5,z:6 --point-source-ez --wavelength 0.004 --time-steps 30",
     (6, 6, 6), (9, 10, 9), 8, (3, 3), (11, 11), 10),
]


def _eps(cfg, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(1.0, 4.0, size=cfg.shape)


@pytest.mark.parametrize("case", CASES)
def test_oracle_adjoint_matches_central_differences(case):
    args, dlo, dhi, mz, mlo, mhi, window = case
    cfg = parse_args(args.split())
    eps = _eps(cfg, 5)
    J, grad = oracle_adjoint(cfg, eps, dlo, dhi, mz, mlo, mhi, window, cfg.time_steps)
    J0 = oracle_objective(cfg, eps, mz, mlo, mhi, window, cfg.time_steps)
    assert J > 0 and abs(J - J0) <= 1e-12 * J
    rng = np.random.default_rng(11)
    shape = grad.shape
    for _ in range(5):
        v = tuple(int(rng.integers(0, shape[a])) for a in range(3))
        p = tuple(dlo[a] + v[a] for a in range(3))
        h = 1e-4 * eps[p]
        ep, em = eps.copy(), eps.copy()
        ep[p] += h
        em[p] -= h
        fd = (oracle_objective(cfg, ep, mz, mlo, mhi, window, cfg.time_steps)
              - oracle_objective(cfg, em, mz, mlo, mhi, window, cfg.time_steps)) / (2 * h)
        scale = np.max(np.abs(grad))
        assert abs(grad[v] - fd) <= 1e-5 * max(abs(fd), 1e-3 * scale), (v, grad[v], fd)
