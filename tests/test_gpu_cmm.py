"""The masked-domain Helmholtz solve on the HIP path (somax_amd.elliptic,
somaxhip_cmm_solve) against the FP64 capacitance-matrix oracle
(oracle/cmm.py), which is pinned by a sparse direct solve.

Bars (relative L2 over the mask): 1e-5 with the plain FP32 inversion, 1e-6
with the mixed one; the FP64 residual of (lap - lam) psi - rhs on the mask
relative to |rhs| at the same levels times the operator's grid-scale
amplification is not asserted, the direct comparison bounds it.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cmm, dst

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _circle(Ny, Nx, r2=0.85):
    yy, xx = np.meshgrid(np.linspace(-1, 1, Ny), np.linspace(-1, 1, Nx), indexing="ij")
    return ((xx**2 + yy**2) < r2).astype(float)


def _octagon_nodes(n):
    c = np.pad(cmm.octagon_center_mask(n, n), 1)
    return ((c[1:, 1:] + c[:-1, 1:] + c[1:, :-1] + c[:-1, :-1]) > 3.5).astype(float)


CASES = [("circle-50x70", lambda: _circle(50, 70)), ("octagon-130", lambda: _octagon_nodes(130)),
         ("octagon-258", lambda: _octagon_nodes(258))]


@pytest.mark.parametrize("name,maskf", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("mixed", 1e-6)])
def test_masked_solve_matches_oracle(gpu, name, maskf, precision, tol):
    from somax_amd.elliptic import MaskedHelmholtz

    mask = maskf()
    dx, dy = 2.0e4, 2.0e4
    f0 = 9.375e-5
    lams = f0**2 * np.array([0.0, 1.0 / 6.0e4**2 / f0**2, 1.0 / 2.0e4**2 / f0**2])  # Poisson and two radii
    _log(f"[cmm] {name} precision={precision}")
    solver = MaskedHelmholtz(mask, dx, dy, lams, precision=precision)
    _log(f"  boundary points: {solver.num_boundary_points}")
    rng = np.random.default_rng(7)
    rhs = rng.standard_normal((3, *mask.shape)).astype(np.float32) * 1e-10
    psi = solver.solve(rhs).cpu().numpy()
    want = cmm.solve(rhs.astype(np.float64), mask, dx, dy, lams, solver.cap, solver.yids, solver.xids)
    rs = [rel(psi[p][mask > 0], want[p][mask > 0]) for p in range(3)]
    _log("  rel-L2 per plane " + " ".join(f"{r:.2e}" for r in rs))
    assert max(rs) < tol
    assert np.all(psi[:, mask == 0] == 0.0)


def test_round_trip(gpu):
    """helmholtz_decomp.ipynb cells 29-33 on the device: f = mask * noise from
    (lap - lam) f * mask."""
    from somax_amd.elliptic import MaskedHelmholtz

    mask = _circle(66, 66)
    lam = [1.0 / 3.0e4**2]
    solver = MaskedHelmholtz(mask, 2e4, 2e4, lam, precision="mixed")
    f = mask * np.random.default_rng(42).standard_normal((1, 66, 66))
    rhs = (dst.helmholtz_apply(f, 2e4, 2e4, lam[0]) * mask).astype(np.float32)
    g = solver.solve(rhs).cpu().numpy()
    r = rel(g[0][mask > 0], f[0][mask > 0])
    _log(f"[cmm] round trip rel-L2 {r:.2e}")
    assert r < 1e-5


def test_rectangle_without_boundary_points_is_the_plain_solve(gpu):
    """A mask of ones has no irregular boundary: the solve is the rectangular
    DST-I inversion."""
    from somax_amd.elliptic import MaskedHelmholtz

    mask = np.ones((34, 34))
    solver = MaskedHelmholtz(mask, 1e4, 1e4, [0.0, 1e-8], precision="mixed")
    assert solver.num_boundary_points == 0
    rhs = np.random.default_rng(1).standard_normal((2, 34, 34)).astype(np.float32)
    psi = solver.solve(rhs).cpu().numpy()
    want = dst.helmholtz_dst(rhs.astype(np.float64), 1e4, 1e4, np.array([0.0, 1e-8]))
    assert rel(psi, want) < 1e-6


def test_errors(gpu):
    from somax_amd.elliptic import MaskedHelmholtz

    with pytest.raises(ValueError):
        MaskedHelmholtz(np.ones((3, 3)), 1.0, 1.0, [0.0])
    solver = MaskedHelmholtz(_circle(20, 20), 1.0, 1.0, [0.0])
    with pytest.raises(ValueError):
        solver.solve(np.zeros((2, 20, 20), np.float32))
