"""GPU checks of peak climbing (pgalf_dev_peak_climb) and of the batched
differentiable pass with climbing (pgalf_batch_run_grad), both additions of
this build (the reference has no gradients).

  * single halo: climbed positions and peak cells bit-identical to the
    oracle's restatement (oracle.peak_climb: the same FP64 gradient
    arithmetic through orc_tsc_grad_d, the same integer ascent); every peak
    is a local maximum of the smoothed grid;
  * batch: forward outputs bit-identical to Batch.run; per halo, gradients
    equal to the single-halo adjoint (FP64 sums in a different order:
    normwise 1e-6) and climbing bit-identical to the single-halo climb --
    the membership of the batched path equals the plain path's;
  * graph replays equal the eager pass (gradients to 1e-6 normwise).
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


@pytest.mark.parametrize("n,ng", [(20000, 48), (8000, 32)])
def test_climb_vs_oracle(n, ng):
    import torch
    pos, mass = halos.small_halo(n, seed=41)
    say("climb n=%d grid=%d" % (n, ng))
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(mass)
    grid = D.grid_struct(ng, ng, ng, 0.0, 0.0, 0.0, 1.0 / ng)
    den = torch.empty(D.padded_shape(ng, ng, ng), dtype=torch.float32, device="cuda")
    D.dev_tsc(x, y, z, m, grid, den)
    D.dev_smooth(den, grid, 2.0 / ng)
    ox, oy, oz, peak = D.dev_peak_climb(x, y, z, grid, den, 5, 0.5)
    D.dev_status()
    got = np.stack([t.cpu().numpy() for t in (ox, oy, oz)], axis=1)
    peak = peak.cpu().numpy()
    F = den.cpu().numpy()
    ref_pos, ref_peak = O.peak_climb(pos, (ng, ng, ng, 0.0, 0.0, 0.0, 1.0 / ng), F, 5, 0.5)
    assert np.array_equal(got.view(np.uint32), ref_pos.view(np.uint32))
    assert np.array_equal(peak, ref_peak)
    Fi = F[..., :ng]
    k, j, i = peak // (ng * ng), (peak // ng) % ng, peak % ng
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                ii, jj, kk = i + dx, j + dy, k + dz
                ok = (ii >= 0) & (ii < ng) & (jj >= 0) & (jj < ng) & (kk >= 0) & (kk < ng)
                assert np.all(Fi[kk[ok], jj[ok], ii[ok]] <= Fi[k[ok], j[ok], i[ok]])
    say("  %d peaks" % np.unique(peak).size)


def _catalogue():
    pos, mass, off, grids, rg = halos.catalogue(config_id=105, nhalo=40, alpha=1.9, nmin=64,
                                                nmax=30000)
    return pos, mass, off, grids, rg


def test_batch_grad_and_climb_equal_single_halo_path():
    import torch
    pos, mass, off, grids, rg = _catalogue()
    n = len(pos)
    say("batch grad+climb: %d halos, %d particles" % (len(off) - 1, n))
    rng = np.random.default_rng(9)
    wt = _dev(rng.normal(size=n).astype(np.float32))
    wk = _dev(rng.normal(size=n).astype(np.float32))
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(mass)
    b = D.Batch(off, grids, rg, K)
    out = b.run_grad(x, y, z, m, wt, wk, rho_mode=2, climb_iters=5, climb_step=0.5)
    b.status()
    o = {k_: v.cpu().numpy() for k_, v in out.items()}
    # forward equals the plain batch forward (unweighted kNN density differs
    # by rho_mode, so compare grid, read-back and lists)
    den = torch.empty(b.grid_floats, dtype=torch.float32, device="cuda")
    rt = torch.empty(n, dtype=torch.float32, device="cuda")
    nb = torch.empty((n, K), dtype=torch.int64, device="cuda")
    rk = torch.empty(n, dtype=torch.float32, device="cuda")
    b.run(x, y, z, m, den, rt, nb, rk)
    b.status()
    assert np.array_equal(o["den"].view(np.uint32), den.cpu().numpy().view(np.uint32))
    assert np.array_equal(o["rho_tsc"].view(np.uint32), rt.cpu().numpy().view(np.uint32))
    assert np.array_equal(o["nbr"], nb.cpu().numpy())
    for h in range(len(off) - 1):
        a, e = int(off[h]), int(off[h + 1])
        g = D.grid_struct(*grids[h])
        xs, ys, zs, ms = x[a:e], y[a:e], z[a:e], m[a:e]
        hden = torch.from_numpy(b.grid_view(o["den"], h).copy()).cuda()
        hnbr = torch.from_numpy(o["nbr"][a:e].copy()).cuda()
        gt = (wt[a:e] / torch.from_numpy(o["rho_tsc"][a:e]).cuda()).contiguous()
        gk = (wk[a:e] / torch.from_numpy(o["rho_knn"][a:e]).cuda()).contiguous()
        sg = D.dev_density_stage_backward(xs, ys, zs, ms, g, float(rg[h]), K, hden, hnbr, 2,
                                          g_tsc=gt, g_knn=gk)
        for name, t in zip(("gx", "gy", "gz", "gm"), sg):
            err = _relerr(o[name][a:e], t.cpu().numpy())
            assert err <= 1e-6, "halo %d %s err %.3g" % (h, name, err)
        cx, cy, cz, pk = D.dev_peak_climb(xs, ys, zs, g, hden, 5, 0.5)
        D.dev_status()
        assert np.array_equal(o["peak"][a:e], pk.cpu().numpy()), "membership, halo %d" % h
        for name, t in zip(("cx", "cy", "cz"), (cx, cy, cz)):
            assert np.array_equal(o[name][a:e].view(np.uint32), t.cpu().numpy().view(np.uint32))
    # graph replays equal the eager pass
    for rep in range(2):
        say("graph replay %d" % rep)
        og = b.run_grad(x, y, z, m, wt, wk, rho_mode=2, climb_iters=5, climb_step=0.5,
                        graph=True, out={k_: torch.empty_like(v) for k_, v in out.items()})
        b.status()
        torch.cuda.synchronize()
        _ = (x * 2.0).sum().item()
        for k_, v in og.items():
            v = v.cpu().numpy()
            if k_ in ("gx", "gy", "gz", "gm"):
                assert _relerr(v, o[k_]) <= 1e-6, k_
            else:
                assert np.array_equal(v.view(np.uint8), o[k_].view(np.uint8)), k_
    b.close()
