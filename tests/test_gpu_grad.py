"""GPU parity of the hand-written adjoint (pgalf_dev_density_stage_backward,
galaxyfinder_amd.autograd.DensityStage) against the oracle's FP64 reference
gradient, which tests/test_oracle_grad.py pins by central finite differences.

L = sum wt log rho_tsc + sum wk log rho_knn with seeded Gaussian weights.
Tolerance: normwise 1e-4 on each of dL/dx, dL/dy, dL/dz, dL/dm (the GPU
forward is FP32 -- grids, densities -- while the reference gradient is FP64
throughout; observed errors are ~1e-6).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from galaxyfinder_amd import density as D  # noqa: E402
from galaxyfinder_amd import halos  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu
K = 32
TOL = 1e-4


def say(msg):
    print("[case] " + msg, flush=True)


def _relerr(got, ref):
    return float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))


def _run(pos, mass, ng, rg_cells, weighted, wt, wk):
    import torch
    from galaxyfinder_amd.autograd import density_stage
    dev = torch.device("cuda")
    x, y, z = (torch.from_numpy(np.ascontiguousarray(pos[:, a])).to(dev).requires_grad_()
               for a in range(3))
    m = torch.from_numpy(np.ascontiguousarray(mass)).to(dev).requires_grad_()
    grid = D.grid_struct(ng, ng, ng, 0.0, 0.0, 0.0, 1.0 / ng)
    rho_t, rho_k, nbr, den = density_stage(x, y, z, m, grid, rg_cells / ng, K, weighted)
    loss = (torch.from_numpy(wt).to(dev) * torch.log(rho_t)).sum() + \
           (torch.from_numpy(wk).to(dev) * torch.log(rho_k)).sum()
    loss.backward()
    D.dev_status()
    g = np.stack([t.grad.cpu().numpy() for t in (x, y, z)], axis=1).astype(np.float64)
    return g, m.grad.cpu().numpy().astype(np.float64), nbr.cpu().numpy(), rho_t, rho_k


@pytest.mark.parametrize("n,ng,weighted", [(20000, 48, True), (30000, 64, False)])
def test_adjoint_vs_fp64_reference_gradient(n, ng, weighted):
    pos, mass = halos.small_halo(n, seed=31)
    mass = (mass * np.random.default_rng(5).uniform(0.5, 1.5, n)).astype(np.float32)
    rng = np.random.default_rng(6)
    wt, wk = rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    say("adjoint n=%d grid=%d weighted=%s" % (n, ng, weighted))
    gp, gm, nbr, _, _ = _run(pos, mass, ng, 2.0, weighted, wt, wk)
    idx, _ = O.knn(pos.astype(np.float64), K)
    assert np.array_equal(idx, nbr), "forward lists differ from the exact kNN"
    grid = (ng, ng, ng, 0.0, 0.0, 0.0, 1.0 / ng)
    rp, rm = O.stage_grad_d(pos.astype(np.float64), mass.astype(np.float64), grid, 2.0 / ng,
                            idx, wt.astype(np.float64), wk.astype(np.float64), weighted)
    for a in range(3):
        e = _relerr(gp[:, a], rp[:, a])
        say("  d/dx%d rel err %.2e" % (a, e))
        assert e <= TOL
    e = _relerr(gm, rm)
    say("  d/dm rel err %.2e" % e)
    assert e <= TOL
    # Euler: both densities are homogeneous of degree 1 in the masses
    lhs = float(np.dot(mass.astype(np.float64), gm))
    # (the unweighted kNN density does not depend on m: only the grid part)
    rhs = float(wt.astype(np.float64).sum() + (wk.astype(np.float64).sum() if weighted else 0.0))
    assert abs(lhs - rhs) <= 1e-4 * (np.abs(wt).sum() + np.abs(wk).sum())


def test_backward_parts_add_up():
    """The two cotangents are independent: backward with both equals the
    sum of the backward with each alone (one adjoint call each)."""
    import torch
    n, ng = 12000, 32
    pos, mass = halos.small_halo(n, seed=32)
    dev = torch.device("cuda")
    x, y, z = (torch.from_numpy(np.ascontiguousarray(pos[:, a])).to(dev) for a in range(3))
    m = torch.from_numpy(mass).to(dev)
    grid = D.grid_struct(ng, ng, ng, 0.0, 0.0, 0.0, 1.0 / ng)
    RG = 2.0 / ng
    den = torch.empty(D.padded_shape(ng, ng, ng), dtype=torch.float32, device=dev)
    rho_t = torch.empty(n, dtype=torch.float32, device=dev)
    nbr = torch.empty((n, K), dtype=torch.int64, device=dev)
    rho_k = torch.empty(n, dtype=torch.float32, device=dev)
    D.dev_tsc(x, y, z, m, grid, den)
    D.dev_smooth(den, grid, RG)
    D.dev_findden(x, y, z, grid, den, rho_t)
    D.dev_knn(x, y, z, K, m=m, nbr=nbr, rho=rho_k, rho_mode=2)
    rng = np.random.default_rng(8)
    gt = torch.from_numpy(rng.normal(size=n).astype(np.float32)).to(dev)
    gk = torch.from_numpy(rng.normal(size=n).astype(np.float32)).to(dev)
    say("backward parts n=%d" % n)
    both = D.dev_density_stage_backward(x, y, z, m, grid, RG, K, den, nbr, 2, g_tsc=gt, g_knn=gk)
    a = D.dev_density_stage_backward(x, y, z, m, grid, RG, K, den, nbr, 2, g_tsc=gt)
    b = D.dev_density_stage_backward(x, y, z, m, grid, RG, K, den, nbr, 2, g_knn=gk)
    D.dev_status()
    for t, u, v in zip(both, a, b):
        t, s = t.cpu().numpy().astype(np.float64), (u + v).cpu().numpy().astype(np.float64)
        assert _relerr(t, s) <= 1e-6
