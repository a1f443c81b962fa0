"""CPU pins of the oracle's FP64 reference gradient of the density stage
(the checker of the HIP adjoint; the reference itself has no adjoint).

L = sum_i wt_i log rho_tsc_i + sum_i wk_i log rho_knn_i with rho_tsc the
read-back of the FFT-smoothed TSC grid and rho_knn the W4 kNN density over
fixed lists.  The analytic gradient must match:
  * central finite differences of the FP64 forward, per part (grid / kNN),
    for positions and masses, weighted and unweighted kNN densities;
  * Euler's identity: both densities are homogeneous of degree 1 in the
    masses, so sum_j m_j dL/dm_j = sum_i (wt_i + wk_i) exactly;
  * translation invariance of the kNN part: sum_j dL/dx_j = 0.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from galaxyfinder_amd import halos  # noqa: E402
from oracle import oracle as O  # noqa: E402

N, NG, K = 2500, 32, 32


@pytest.fixture(scope="module")
def problem():
    pos, mass = halos.small_halo(N, seed=11)
    pos = pos.astype(np.float64)
    # unequal masses so the mass gradient is not degenerate
    mass = mass.astype(np.float64) * np.random.default_rng(3).uniform(0.5, 1.5, N)
    grid = (NG, NG, NG, 0.0, 0.0, 0.0, 1.0 / NG)
    RG = 2.0 / NG
    idx, _ = O.knn(pos, K)
    rng = np.random.default_rng(4)
    return pos, mass, grid, RG, idx, rng.normal(size=N), rng.normal(size=N)


def _fd_pos(pos, mass, grid, RG, idx, wt, wk, weighted, i, a, h=1e-6):
    p1, p2 = pos.copy(), pos.copy()
    p1[i, a] += h
    p2[i, a] -= h
    return (O.stage_loss_d(p1, mass, grid, RG, idx, wt, wk, weighted) -
            O.stage_loss_d(p2, mass, grid, RG, idx, wt, wk, weighted)) / (2 * h)


def _fd_mass(pos, mass, grid, RG, idx, wt, wk, weighted, i, rel=1e-6):
    h = rel * mass[i]
    m1, m2 = mass.copy(), mass.copy()
    m1[i] += h
    m2[i] -= h
    return (O.stage_loss_d(pos, m1, grid, RG, idx, wt, wk, weighted) -
            O.stage_loss_d(pos, m2, grid, RG, idx, wt, wk, weighted)) / (2 * h)


@pytest.mark.parametrize("part", ["tsc", "knn_weighted", "knn_unweighted"])
def test_gradient_matches_finite_differences(problem, part):
    pos, mass, grid, RG, idx, wt, wk = problem
    zero = np.zeros(N)
    weighted = part != "knn_unweighted"
    wt_, wk_ = (wt, zero) if part == "tsc" else (zero, wk)
    gp, gm = O.stage_grad_d(pos, mass, grid, RG, idx, wt_, wk_, weighted)
    sel = np.random.default_rng(7).choice(N, 6, replace=False)
    for i in sel:
        for a in range(3):
            fd = _fd_pos(pos, mass, grid, RG, idx, wt_, wk_, weighted, i, a)
            assert abs(fd - gp[i, a]) <= 1e-6 * max(abs(gp[i, a]), 1.0), (i, a, fd, gp[i, a])
        if part != "knn_unweighted":
            fd = _fd_mass(pos, mass, grid, RG, idx, wt_, wk_, weighted, i)
            assert abs(fd - gm[i]) <= 1e-5 * max(abs(gm[i]), 1.0), (i, fd, gm[i])
    if part == "knn_unweighted":
        assert np.all(gm == 0.0)


def test_neighbour_terms_reach_the_neighbours(problem):
    """Moving a particle changes its neighbours' kNN densities: the FD of a
    loss that weights only particle i's density must see particle j != i."""
    pos, mass, grid, RG, idx, _, _ = problem
    i = 100
    wk = np.zeros(N)
    wk[i] = 1.0
    gp, _ = O.stage_grad_d(pos, mass, grid, RG, idx, np.zeros(N), wk, True)
    j = idx[i, 5]
    assert j != i and np.abs(gp[j]).max() > 0
    for a in range(3):
        fd = _fd_pos(pos, mass, grid, RG, idx, np.zeros(N), wk, True, j, a)
        assert abs(fd - gp[j, a]) <= 1e-6 * max(abs(gp[j, a]), 1.0)
    far = np.setdiff1d(np.arange(N), np.append(idx[i], i))
    assert np.all(gp[far] == 0.0)


def test_euler_mass_identity_and_translation(problem):
    pos, mass, grid, RG, idx, wt, wk = problem
    gp, gm = O.stage_grad_d(pos, mass, grid, RG, idx, wt, wk, True)
    lhs = float(np.dot(mass, gm))
    rhs = float(wt.sum() + wk.sum())
    assert abs(lhs - rhs) <= 1e-9 * max(1.0, np.abs(wt).sum() + np.abs(wk).sum())
    gk, _ = O.stage_grad_d(pos, mass, grid, RG, idx, np.zeros(N), wk, True)
    assert np.abs(gk.sum(axis=0)).max() <= 1e-9 * np.abs(gk).sum()


def test_peak_climb_restatement_properties():
    """The oracle's climb (the GPU climb's checker): peaks are local maxima
    of the smoothed grid; with no ascent steps the peak is reached from the
    nearest node; positions stay within [1, n-2] cells."""
    pos, mass = halos.small_halo(3000, seed=12)
    ng = 32
    grid = (ng, ng, ng, 0.0, 0.0, 0.0, 1.0 / ng)
    den, _ = O.tsc_deposit(pos, mass, grid)
    sm = O.gsmooth(den, ng, ng, ng, 1.0 / ng, 2.0 / ng)
    F = sm[..., :ng]
    for niter in (0, 5):
        out, peak = O.peak_climb(pos, grid, sm, niter, 0.5)
        q = out.astype(np.float64) * ng
        assert q.min() >= 1.0 - 1e-5 and q.max() <= ng - 2 + 1e-5
        k, j, i = peak // (ng * ng), (peak // ng) % ng, peak % ng
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    ii, jj, kk = i + dx, j + dy, k + dz
                    ok = (ii >= 0) & (ii < ng) & (jj >= 0) & (jj < ng) & (kk >= 0) & (kk < ng)
                    assert np.all(F[kk[ok], jj[ok], ii[ok]] <= F[k[ok], j[ok], i[ok]])
    out0, _ = O.peak_climb(pos, grid, sm, 0, 0.5)
    assert np.allclose(out0, pos, atol=1e-6)
