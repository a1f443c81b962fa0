"""GPU parity of the batched catalogue path (pgalf_batch_* through ctypes).

Every halo of a batch must come out bit-identical to the single-halo device
path on the same particles -- deposit (integer fixed-point sums with the
halo's own scale), smoothing, read-back, kNN lists and W4 densities -- which
the single-halo parity tests pin against the oracle and the reference
fixtures.  A few halos are also checked directly against the oracle, and
graph replays against the eager run.
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


def _catalogue():
    """60 halos of 100..30000 particles (power law), grids 32..128 from n, a
    few non-power-of-two grids and two smoothing lengths."""
    pos, mass, off, grids, rg = halos.catalogue(config_id=103, nhalo=60, nmax=30000)
    grids = list(grids)
    for h in (3, 17, 40):  # non-power-of-two grids (padded-FFT smoothing path)
        n = 40 + 2 * (h % 3)
        grids[h] = (n, n, n, 0.0, 0.0, 0.0, 1.0 / n)
        rg[h] = 1.5 / n
    return pos, mass, off, grids, rg


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _run_batch(b, pos, mass, graph=False):
    import torch
    n = len(pos)
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(mass)
    den = torch.full((b.grid_floats,), -1.0, dtype=torch.float32, device="cuda")
    rho_t = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    nbr = torch.full((n, K), -1, dtype=torch.int64, device="cuda")
    rho_k = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    b.run(x, y, z, m, den, rho_t, nbr, rho_k, graph=graph)
    b.status()
    return (x, y, z, m, den, rho_t, nbr, rho_k)


def _single(pos, mass, grid, RG):
    import torch
    n = len(pos)
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(mass)
    g = D.grid_struct(*grid)
    den = torch.empty(D.padded_shape(grid[0], grid[1], grid[2]), dtype=torch.float32,
                      device="cuda")
    rho_t = torch.empty(n, dtype=torch.float32, device="cuda")
    nbr = torch.empty((n, K), dtype=torch.int64, device="cuda")
    rho_k = torch.empty(n, dtype=torch.float32, device="cuda")
    D.dev_density_stage(x, y, z, m, g, float(RG), K, den, rho_t, nbr, rho_k)
    D.dev_status()
    return den.cpu().numpy(), rho_t.cpu().numpy(), nbr.cpu().numpy(), rho_k.cpu().numpy()


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_batch_equals_single_halo_path():
    pos, mass, off, grids, rg = _catalogue()
    say("batch of %d halos, %d particles" % (len(off) - 1, off[-1]))
    b = D.Batch(off, grids, rg, K)
    out = _run_batch(b, pos, mass)
    den, rho_t, nbr, rho_k = (t.cpu().numpy() for t in out[4:])
    for h in range(len(off) - 1):
        a, e = off[h], off[h + 1]
        say("halo %d: n=%d grid=%d" % (h, e - a, grids[h][0]))
        sden, srt, snb, srk = _single(pos[a:e], mass[a:e], grids[h], rg[h])
        assert np.array_equal(_bits(b.grid_view(den, h)), _bits(sden)), "grid, halo %d" % h
        assert np.array_equal(_bits(rho_t[a:e]), _bits(srt)), "read-back, halo %d" % h
        assert np.array_equal(nbr[a:e], snb), "kNN lists, halo %d" % h
        assert np.array_equal(_bits(rho_k[a:e]), _bits(srk)), "W4 density, halo %d" % h
    b.close()


def test_batch_halos_vs_oracle():
    pos, mass, off, grids, rg = _catalogue()
    b = D.Batch(off, grids, rg, K)
    out = _run_batch(b, pos, mass)
    den, rho_t, nbr, rho_k = (t.cpu().numpy() for t in out[4:])
    for h in (0, 3, 17, int(np.argmax(np.diff(off)))):
        a, e = off[h], off[h + 1]
        p, m = pos[a:e].astype(np.float64), mass[a:e].astype(np.float64)
        say("oracle halo %d n=%d grid=%d" % (h, e - a, grids[h][0]))
        g = grids[h]
        dep, bad = O.tsc_deposit(p, m, g, fp64=True)
        assert bad == 0
        nx = g[0]
        ref = O.gsmooth_f64(np.ascontiguousarray(dep[..., :nx]), g[6], float(rg[h]))
        got = b.grid_view(den, h)[..., :nx].astype(np.float64)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 1e-6, "smoothed grid error %.3g" % err
        rb, _ = O.findden(p, g, b.grid_view(den, h))
        assert np.array_equal(_bits(rb), _bits(rho_t[a:e]))
        idx, d2 = O.knn(p, K)
        assert np.array_equal(idx, nbr[a:e])
        assert np.array_equal(_bits(O.sph_density(idx, d2)), _bits(rho_k[a:e]))
    b.close()


def test_batch_graph_replays_match_eager():
    import torch
    pos, mass, off, grids, rg = _catalogue()
    b = D.Batch(off, grids, rg, K)
    ref = [t.cpu().numpy() for t in _run_batch(b, pos, mass)[4:]]
    x, y, z, m = _dev(pos[:, 0]), _dev(pos[:, 1]), _dev(pos[:, 2]), _dev(mass)
    n = len(pos)
    den = torch.empty(b.grid_floats, dtype=torch.float32, device="cuda")
    rho_t = torch.empty(n, dtype=torch.float32, device="cuda")
    nbr = torch.empty((n, K), dtype=torch.int64, device="cuda")
    rho_k = torch.empty(n, dtype=torch.float32, device="cuda")
    for rep in range(3):
        say("graph replay %d" % rep)
        for t in (den, rho_t, rho_k):
            t.fill_(-7.0)
        nbr.fill_(-7)
        torch.cuda.synchronize()
        b.run(x, y, z, m, den, rho_t, nbr, rho_k, graph=True)
        b.status()
        got = [t.cpu().numpy() for t in (den, rho_t, nbr, rho_k)]
        for g_, r_ in zip(got, ref):
            assert np.array_equal(g_.view(np.uint8), r_.view(np.uint8)), "replay %d" % rep
        # eager work on the same stream between replays
        torch.cuda.synchronize()
        _ = (x * 2.0).sum().item()
    b.close()
