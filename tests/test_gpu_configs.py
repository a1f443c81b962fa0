"""GPU checks at the two workloads that exercise the adjoint and the batched
differentiable pass at their own sizes:

  * one halo of 1,048,576 particles (NFW c=8 + 10 subhalos), 128^3 grid,
    RG = 2 cells, k = 32, mass-weighted W4, L = sum wt log rho_tsc +
    sum wk log rho_knn with seeded Gaussian weights:
      - every neighbour list bit-exact against the oracle's exact kNN;
      - dL/dx, dL/dy, dL/dz, dL/dm against the oracle's FP64 analytic
        gradient (oracle.stage_grad_d, pinned by central finite differences in
        tests/test_oracle_grad.py): normwise over the halo, and pointwise on
        256 seeded particles away from the TSC weight kinks (|d| within 1e-4
        box units of 0.5 or 1.5 cells on any axis);
      - Euler's identity: both densities are homogeneous of degree 1 in the
        masses, so sum m dL/dm = sum wt + sum wk;
  * 50,000 halos (dN/dn ~ n^-1.9 on [64, 2097152], grids 32^3-256^3, RG = 2
    cells, k = 32, mass-weighted W4, 5 climbing steps of 0.5 cell) through
    one batched forward + backward + climbing pass, eager and from a HIP
    graph:
      - membership (each particle's peak cell) of the 4 largest and 20
        seeded halos bit-identical to the single-halo climb;
      - finite gradients; Euler's identity per halo;
      - the batch's device scratch stays well inside the device memory.
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


def say(msg):
    print("[case] " + msg, flush=True)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _relerr(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _near_kink(pos, grid, h_fd=1e-4):
    """Particles whose TSC offset d (in cells, from the nearest node) lies
    within h_fd box units of a weight kink (|d| = 0.5 or 1.5) on any axis."""
    nx, ny, nz, x0, y0, z0, h = grid
    xp = (pos - np.array([x0, y0, z0])) / h
    d = np.abs(xp - np.rint(xp))
    tol = h_fd / h
    return ((np.abs(d - 0.5) < tol) | (np.abs(d - 1.5) < tol)).any(axis=1)


def test_differentiation_workload_1m_particles():
    import torch
    from galaxyfinder_amd.autograd import density_stage
    pos, mass, spec = halos.single_halo_config(4)
    n, ng = len(pos), spec["nx"]
    grid = (ng, ng, ng, spec["xmin"], spec["ymin"], spec["zmin"], spec["cellsize"])
    rng = np.random.default_rng(20261203 + 4)
    wt = rng.normal(size=n).astype(np.float32)
    wk = rng.normal(size=n).astype(np.float32)
    say("C4 adjoint: N=%d grid=%d^3 RG=%.1f cells" % (n, ng, spec["rg_cells"]))
    leaves = [_dev(pos[:, a]).requires_grad_() for a in range(3)] + [_dev(mass).requires_grad_()]
    g = D.grid_struct(*grid)
    rho_t, rho_k, nbr, _ = density_stage(leaves[0], leaves[1], leaves[2], leaves[3], g,
                                         spec["RG"], K, True)
    loss = (_dev(wt) * torch.log(rho_t)).sum() + (_dev(wk) * torch.log(rho_k)).sum()
    loss.backward()
    D.dev_status()
    gp = np.stack([t.grad.cpu().numpy() for t in leaves[:3]], axis=1).astype(np.float64)
    gm = leaves[3].grad.cpu().numpy().astype(np.float64)
    nbr = nbr.cpu().numpy()
    p64 = pos.astype(np.float64)
    idx, _ = O.knn(p64, K)
    assert np.array_equal(idx, nbr), "neighbour lists differ from the exact kNN"
    rp, rm = O.stage_grad_d(p64, mass.astype(np.float64), grid, spec["RG"], idx,
                            wt.astype(np.float64), wk.astype(np.float64), True)
    errs = [_relerr(gp[:, a], rp[:, a]) for a in range(3)] + [_relerr(gm, rm)]
    say("  normwise rel err dx %.2e dy %.2e dz %.2e dm %.2e" % tuple(errs))
    assert max(errs) <= 1e-4
    # 256 seeded particles away from the kinks, pointwise against the largest
    # component of the particle's reference gradient
    kink = _near_kink(p64, grid)
    pick = np.random.default_rng(4).choice(np.flatnonzero(~kink), 256, replace=False)
    scale = np.maximum(np.abs(rp[pick]).max(axis=1), 1e-300)
    pt = (np.abs(gp[pick] - rp[pick]).max(axis=1) / scale).max()
    ptm = (np.abs(gm[pick] - rm[pick]) / np.maximum(np.abs(rm[pick]), 1e-300)).max()
    say("  256 particles (%d near kinks excluded overall): max pointwise %.2e (x), %.2e (m)"
        % (int(kink.sum()), pt, ptm))
    assert pt <= 1e-3 and ptm <= 1e-3
    lhs = float(np.dot(mass.astype(np.float64), gm))
    rhs = float(wt.astype(np.float64).sum() + wk.astype(np.float64).sum())
    say("  Euler: sum m dL/dm = %.6f, sum w = %.6f" % (lhs, rhs))
    assert abs(lhs - rhs) <= 1e-4 * (np.abs(wt).sum() + np.abs(wk).sum())


def test_batched_climbing_workload_50k_halos():
    import torch
    say("C5: generating the 50000-halo catalogue")
    pos, mass, off, grids, rg = halos.catalogue_vec(5)
    n, nh = len(pos), len(off) - 1
    say("C5: %d halos, %d particles" % (nh, n))
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(mass)
    del pos
    rng = np.random.default_rng(20261203 + 5)
    wt = rng.normal(size=n).astype(np.float32)
    wk = rng.normal(size=n).astype(np.float32)
    b = D.Batch(off, grids, rg, K)
    out = b.run_grad(x, y, z, m, _dev(wt), _dev(wk), rho_mode=2, climb_iters=5, climb_step=0.5)
    b.status()
    say("C5: eager pass done; scratch peak %.2f GB" % (b.workspace_bytes() / 1e9))
    total = torch.cuda.get_device_properties(0).total_memory
    assert b.workspace_bytes() < 0.5 * total
    peak = out["peak"].cpu().numpy()
    gm = out["gm"].cpu().numpy().astype(np.float64)
    for name in ("gx", "gy", "gz"):
        assert bool(torch.isfinite(out[name]).all()), name
    assert np.isfinite(gm).all()
    # per-halo Euler identity: sum over the halo of m dL/dm = sum of its weights
    lhs = np.add.reduceat(mass.astype(np.float64) * gm, off[:-1])
    rhs = np.add.reduceat(wt.astype(np.float64) + wk.astype(np.float64), off[:-1])
    scale = np.add.reduceat(np.abs(wt).astype(np.float64) + np.abs(wk), off[:-1])
    bad = np.abs(lhs - rhs) > 1e-4 * scale
    say("C5: Euler identity, worst %.2e of the weight scale" % float(np.max(np.abs(lhs - rhs) / scale)))
    assert not bad.any(), "Euler identity fails on %d halos" % int(bad.sum())
    # graph replay equals the eager pass
    og = b.run_grad(x, y, z, m, _dev(wt), _dev(wk), rho_mode=2, climb_iters=5, climb_step=0.5,
                    graph=True)
    b.status()
    assert np.array_equal(og["peak"].cpu().numpy(), peak)
    assert np.array_equal(og["nbr"].cpu().numpy().view(np.uint8)[:1 << 20],
                          out["nbr"].cpu().numpy().view(np.uint8)[:1 << 20])
    # membership against the single-halo climb
    counts = np.diff(off)
    pick = list(np.argsort(counts)[-4:]) + list(
        np.random.default_rng(5).choice(nh, 20, replace=False))
    for h in pick:
        a, e = int(off[h]), int(off[h + 1])
        g = D.grid_struct(*grids[h])
        den = torch.empty(D.padded_shape(grids[h][0], grids[h][1], grids[h][2]),
                          dtype=torch.float32, device="cuda")
        D.dev_tsc(x[a:e], y[a:e], z[a:e], m[a:e], g, den)
        D.dev_smooth(den, g, float(rg[h]))
        _, _, _, pk = D.dev_peak_climb(x[a:e], y[a:e], z[a:e], g, den, 5, 0.5)
        D.dev_status()
        assert np.array_equal(pk.cpu().numpy(), peak[a:e]), "membership, halo %d" % h
    say("C5: membership identical on %d halos" % len(pick))
    b.close()


def test_calls_on_two_streams_share_the_workspace_safely():
    """Two stage calls on different streams reuse one device arena: each gets
    the same outputs as the same call alone on the default stream."""
    import torch
    outs = []
    for seed, n, ng in ((51, 60000, 64), (52, 40000, 48)):
        pos, mass = halos.small_halo(n, seed=seed)
        outs.append((pos, mass, n, ng))

    def run(pos, mass, n, ng):
        x, y, z = (_dev(pos[:, a]) for a in range(3))
        m = _dev(mass)
        g = D.grid_struct(ng, ng, ng, 0.0, 0.0, 0.0, 1.0 / ng)
        den = torch.empty(D.padded_shape(ng, ng, ng), dtype=torch.float32, device="cuda")
        rt = torch.empty(n, dtype=torch.float32, device="cuda")
        nb = torch.empty((n, K), dtype=torch.int64, device="cuda")
        rk = torch.empty(n, dtype=torch.float32, device="cuda")
        D.dev_density_stage(x, y, z, m, g, 2.0 / ng, K, den, rt, nb, rk)
        return den, rt, nb, rk

    ref = []
    for o in outs:
        r = run(*o)
        D.dev_status()
        ref.append([t.cpu().numpy() for t in r])
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    got = []
    for o, st in zip(outs, (s1, s2)):
        with torch.cuda.stream(st):
            got.append(run(*o))
    torch.cuda.synchronize()
    D.dev_status()
    for r, g_ in zip(ref, got):
        for a, b_ in zip(r, g_):
            assert np.array_equal(a.view(np.uint8), b_.cpu().numpy().view(np.uint8))
