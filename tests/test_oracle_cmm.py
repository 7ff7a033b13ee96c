"""The capacitance-matrix restatement (oracle/cmm.py) pinned by an independent
sparse direct solve, by the round trip of helmholtz_decomp.ipynb (cells
29-33) and by the octagon homogeneous-solution means that
mqg_demo_octagonal.ipynb prints (cell 12); plus the HIP path's host set-up
(somax_amd.elliptic) against the oracle's."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cmm, dst, modes


def _circle(Ny, Nx, r2=0.85):
    yy, xx = np.meshgrid(np.linspace(-1, 1, Ny), np.linspace(-1, 1, Nx), indexing="ij")
    return ((xx**2 + yy**2) < r2).astype(float)


def _octagon_nodes(n):
    """Node mask of the octagon: a node is inside when its four cells are."""
    c = np.pad(cmm.octagon_center_mask(n, n), 1)
    return ((c[1:, 1:] + c[:-1, 1:] + c[1:, :-1] + c[:-1, :-1]) > 3.5).astype(float)


@pytest.mark.parametrize("maskf", [lambda: _circle(34, 42), lambda: _octagon_nodes(40)])
def test_cmm_equals_sparse_direct(maskf):
    mask = maskf()
    yb, xb = cmm.irregular_boundary(mask)
    dx, dy = 3e4, 2.5e4
    lams = np.array([0.0, 2e-10, 1e-9])
    cap = cmm.capacitance_matrices(mask.shape, dx, dy, lams, yb, xb)
    r = np.random.default_rng(0).standard_normal((3, *mask.shape))
    psi = cmm.solve(r, mask, dx, dy, lams, cap, yb, xb)
    for p, lam in enumerate(lams):
        ref = cmm.masked_direct(r[p], mask, dx, dy, lam)
        assert np.abs(psi[p] - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.all(psi[:, mask == 0] == 0.0)


def test_cmm_round_trip():
    """helmholtz_decomp.ipynb cells 29-33: f = mask * noise is recovered from
    (lap - 1) f * mask (lam = 1 in grid units there)."""
    mask = _circle(40, 40)
    yb, xb = cmm.irregular_boundary(mask)
    f = mask * np.random.default_rng(42).standard_normal((1, 40, 40))
    rhs = dst.helmholtz_apply(f, 1.0, 1.0, 1.0) * mask
    cap = cmm.capacitance_matrices(mask.shape, 1.0, 1.0, [1.0], yb, xb)
    g = cmm.solve(rhs, mask, 1.0, 1.0, [1.0], cap, yb, xb)
    assert np.abs(g - f).max() < 1e-12


def test_boundary_is_the_notebook_rule():
    """A point is on the irregular boundary iff it is outside the mask and its
    3x3 window (cells 23-25, avg_pool2d > 0) meets the mask."""
    mask = _circle(20, 24)
    yb, xb = cmm.irregular_boundary(mask)
    got = set(zip(yb.tolist(), xb.tolist()))
    want = set()
    for j in range(20):
        for i in range(24):
            win = mask[max(j - 1, 0):j + 2, max(i - 1, 0):i + 2]
            if mask[j, i] == 0 and win.sum() > 0:
                want.add((j, i))
    assert got == want


def test_octagon_homsol_known_answers():
    """mqg_demo_octagonal.ipynb cells 4, 6, 12 (x64): 256 nodes, L = 5120 km,
    dx = L/(N-1), f0 = 9.375e-5, rigid-lid modes of H = (400, 1100, 2600),
    g' = (0.025, 0.0125); the rectangle's homogeneous solution, centre-averaged
    and multiplied by the octagon cell mask.  Printed: HOMSOL mean 3.565129e-01;
    HOMSOL MEAN min 1.083716e-02, mean 2.999800e-01, max 8.720492e-01."""
    N, L, f0 = 256, 5120e3, 9.375e-5
    d = L / (N - 1)
    A = modes.coupling_matrix([400.0, 1100.0, 2600.0], [9.81, 0.025, 0.0125], free_surface=False)
    lam = f0**2 * np.sort(np.linalg.eigvals(A).real)
    hs = []
    for lm in lam:
        h = np.ones((N, N))
        h[1:-1, 1:-1] += dst.helmholtz_dst(lm * np.ones((N - 2, N - 2)), d, d, lm)
        hs.append(h)
    hs = np.array(hs)
    assert hs.mean() == pytest.approx(3.565129e-01, rel=2e-5)
    cell = 0.25 * (hs[:, 1:, 1:] + hs[:, :-1, 1:] + hs[:, 1:, :-1] + hs[:, :-1, :-1]) * cmm.octagon_center_mask(N, N)
    means = cell.mean(axis=(1, 2))
    assert means.min() == pytest.approx(1.083716e-02, rel=2e-5)
    assert means.mean() == pytest.approx(2.999800e-01, rel=2e-5)
    assert means.max() == pytest.approx(8.720492e-01, rel=2e-5)


def test_host_setup_matches_oracle():
    from somax_amd import elliptic

    mask = _octagon_nodes(36)
    yb, xb = cmm.irregular_boundary(mask)
    y2, x2 = elliptic.irregular_boundary(mask)
    assert np.array_equal(yb, y2) and np.array_equal(xb, x2)
    lams = [0.0, 4e-10]
    want = cmm.capacitance_matrices(mask.shape, 2e4, 3e4, lams, yb, xb)
    got = elliptic.compute_capacitance_matrices(mask.shape, 2e4, 3e4, lams, yb, xb, batch=7)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
