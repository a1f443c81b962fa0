"""GPU edge cases of the density stage.

The cases:
  * neighbour counts below 32;
  * particle counts that leave a partial last bucket of 8, down to exactly k
    particles;
  * coincident particles (distance ties, broken by index);
  * too few candidates (status 5, where the reference exits);
  * empty deposits and read-backs;
  * a batch of tiny ragged halos;
  * a device grid that is not 8-byte aligned (refused);
  * queries with two neighbours at exactly equal distance, spread so that
    each wave of the packed-key query flags at most a few lanes: those
    queries are re-run alone by the radius kernel (the rest of the wave
    keeps the packed results).

Expected results:
  * lists come from the oracle's brute-force exact search and are bit-exact;
  * W4 densities are bit-exact against the oracle's restatement of
    nnost.c:607-623;
  * every halo of the batch is bit-identical to the single-halo path.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from galaxyfinder_amd import density as D  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print("[case] " + msg, flush=True)


def _cloud(n, k, seed):
    """n points in [0.3, 0.7)^3 with min(k - 1, n // 10) of them coincident
    (ties at distance 0, yet every k-th neighbour distance stays positive)."""
    rng = np.random.default_rng(seed)
    pos = 0.3 + 0.4 * rng.random((n, 3))
    c = max(0, min(k - 1, n // 10))
    pos[1:c] = pos[0]
    return pos


@pytest.mark.parametrize("n,k", [(32, 32), (33, 32), (9, 9), (100, 2), (257, 17), (1001, 8),
                                 (4099, 32), (100, 1)])
def test_knn_ragged_sizes_and_small_k(n, k):
    say("findsphdensity n=%d k=%d" % (n, k))
    pos = _cloud(n, k, 1000 + n + k)
    bp = D.make_particles(pos)
    nbr = np.zeros((n, k), dtype=np.int64)
    rho = np.zeros(n, dtype=np.float32)
    D.findsphdensity(bp, n, nbr, k, rho)
    idx, d2 = O.knn(pos, k, method="brute")
    assert np.array_equal(nbr, idx)
    if k > 1:  # k = 1: the neighbour distance is the query's own, 0 (W4 undefined)
        assert np.array_equal(rho.view(np.int32), O.sph_density(idx, d2).view(np.int32))
    nl = np.zeros((n, k), dtype=np.int64)
    D.near_lists(bp, n, k, nl)
    assert np.array_equal(nl, idx)


def test_short_list_is_reported():
    say("short list n=20 k=32")
    bp = D.make_particles(_cloud(20, 32, 3))
    with pytest.raises(D.PgalfError) as ei:
        D.findsphdensity(bp, 20, np.zeros((20, 32), dtype=np.int64), 32,
                         np.zeros(20, dtype=np.float32))
    assert ei.value.code == 5
    with pytest.raises(D.PgalfError) as ei:
        D.near_lists(bp, 20, 32, np.zeros((20, 32), dtype=np.int64))
    assert ei.value.code == 5


def test_empty_deposit_and_readback():
    say("empty deposit / read-back")
    bp = D.make_particles(np.zeros((0, 3)))
    den = np.full(D.padded_shape(16, 16, 16), 7.0, dtype=np.float32)
    D.assign_density_TSC(bp, 0, den, 16, 16, 16, 0.0, 0.0, 0.0, 1.0 / 16, 5)
    assert np.all(den == 0.0)
    out = np.zeros(0, dtype=np.float32)
    D.findDen(bp, 0, out, 0.0, 0.0, 0.0, 1.0 / 16, den, 16, 16, 16)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("k", [32, 8])
def test_batch_of_tiny_ragged_halos(k):
    """Halos of exactly k, k + 1 and other non-multiple-of-8 sizes, on 16^3,
    20^3 and 32^3 grids, against the single-halo path."""
    import torch
    sizes = [k, k + 1, 40, 47, 64, 65, 100, 257, 1000, 33 + k]
    say("batch of %d tiny halos, k=%d" % (len(sizes), k))
    pos = np.concatenate([_cloud(n, k, 77 + i) for i, n in enumerate(sizes)])
    rng = np.random.default_rng(5)
    mass = (0.5 + rng.random(len(pos))).astype(np.float32)
    pos = pos.astype(np.float32)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    grids, rg = [], []
    for i in range(len(sizes)):
        n = (16, 20, 32)[i % 3]
        grids.append((n, n, n, 0.0, 0.0, 0.0, 1.0 / n))
        rg.append(1.5 / n)
    b = D.Batch(off, grids, np.array(rg), k)
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(mass)
    npart = len(pos)
    den = torch.full((b.grid_floats,), -1.0, dtype=torch.float32, device="cuda")
    rho_t = torch.full((npart,), -1.0, dtype=torch.float32, device="cuda")
    nbr = torch.full((npart, k), -1, dtype=torch.int64, device="cuda")
    rho_k = torch.full((npart,), -1.0, dtype=torch.float32, device="cuda")
    b.run(x, y, z, m, den, rho_t, nbr, rho_k)
    b.status()
    den_h, rho_t_h = den.cpu().numpy(), rho_t.cpu().numpy()
    nbr_h, rho_k_h = nbr.cpu().numpy(), rho_k.cpu().numpy()
    for h, n in enumerate(sizes):
        a, e = int(off[h]), int(off[h + 1])
        g = D.grid_struct(*grids[h])
        sden = torch.empty(D.padded_shape(*grids[h][:3]), dtype=torch.float32, device="cuda")
        srt = torch.empty(n, dtype=torch.float32, device="cuda")
        snb = torch.empty((n, k), dtype=torch.int64, device="cuda")
        srk = torch.empty(n, dtype=torch.float32, device="cuda")
        D.dev_density_stage(x[a:e], y[a:e], z[a:e], m[a:e], g, float(rg[h]), k, sden, srt,
                            snb, srk)
        D.dev_status()
        assert np.array_equal(b.grid_view(den_h, h).view(np.int32),
                              sden.cpu().numpy().view(np.int32)), "grid, halo %d" % h
        assert np.array_equal(rho_t_h[a:e].view(np.int32), srt.cpu().numpy().view(np.int32))
        assert np.array_equal(nbr_h[a:e], snb.cpu().numpy()), "lists, halo %d" % h
        assert np.array_equal(rho_k_h[a:e].view(np.int32), srk.cpu().numpy().view(np.int32))
        # and the lists against the oracle's exact search
        idx, _ = O.knn(pos[a:e].astype(np.float64), k, method="brute")
        assert np.array_equal(nbr_h[a:e], idx), "oracle lists, halo %d" % h
    b.close()


def test_unaligned_grid_is_rejected():
    """Grids are written as float pairs: a device grid at an odd float offset
    is refused before any launch (PGALF_ERR_ARG)."""
    import torch
    say("unaligned device grid")
    pos = _cloud(64, 32, 9).astype(np.float32)
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(np.ones(64, dtype=np.float32))
    g = D.grid_struct(16, 16, 16, 0.0, 0.0, 0.0, 1.0 / 16)
    nfl = int(np.prod(D.padded_shape(16, 16, 16)))
    big = torch.zeros(nfl + 1, dtype=torch.float32, device="cuda")
    with pytest.raises(D.PgalfError) as ei:
        D.dev_tsc(x, y, z, m, g, big[1:].view(D.padded_shape(16, 16, 16)))
    assert ei.value.code == 1
    D.dev_tsc(x, y, z, m, g, big[:nfl].view(D.padded_shape(16, 16, 16)))
    D.dev_status()
    unit, bad = O.tsc_deposit(np.array([[0.5, 0.5, 0.5]]), np.array([1.0]),
                              (16, 16, 16, 0.0, 0.0, 0.0, 1.0 / 16), fp64=True)
    assert bad == 0
    tot, want = float(big[:nfl].double().sum()), 64.0 * float(unit.sum())
    assert abs(tot - want) <= 1e-5 * want


@pytest.mark.parametrize("k,rho_mode", [(32, 1), (20, 2)])
def test_single_flagged_queries_rerun_by_radius(k, rho_mode):
    """20,000 float32 points in [0.1, 0.9)^3, plus, for 60 queries q on a
    coarse grid of the box, the two points q +- v (v a multiple of 2^-14 per
    axis, so both are exact in float32 and exactly equidistant from q): each
    such q has two of its first k at the same distance, which the packed
    keys cannot order, so its lane is flagged (and, the queries lying far
    apart, few lanes per wave).  Every list and density against the exact
    search of the oracle."""
    import torch
    rng = np.random.default_rng(77 + k)
    base = (0.1 + 0.8 * rng.random((20000, 3))).astype(np.float32)
    g = np.linspace(0.15, 0.85, 4)
    targets = np.stack(np.meshgrid(g, g, g[:4]), -1).reshape(-1, 3)[:60]
    picks = [int(np.argmin(((base - t) ** 2).sum(1))) for t in targets]
    v = np.array([3.0, -2.0, 1.0], dtype=np.float32) * np.float32(2.0 ** -14)
    extra = np.concatenate([np.stack([base[i] + v, base[i] - v]) for i in picks])
    pos = np.concatenate([base, extra]).astype(np.float32)
    # the planted pairs are exact and exactly equidistant from their query
    for j, i in enumerate(picks):
        a, b = pos[20000 + 2 * j].astype(np.float64), pos[20001 + 2 * j].astype(np.float64)
        q = pos[i].astype(np.float64)
        assert ((q - a) ** 2).sum() == ((q - b) ** 2).sum()
    n = len(pos)
    say("planted ties: %d queries, n=%d k=%d rho_mode=%d" % (len(picks), n, k, rho_mode))
    mass = (0.5 + rng.random(n)).astype(np.float32)
    x, y, z = (torch.from_numpy(np.ascontiguousarray(pos[:, a])).cuda() for a in range(3))
    m = torch.from_numpy(mass).cuda()
    nbr = torch.empty((n, k), dtype=torch.int64, device="cuda")
    rho = torch.empty(n, dtype=torch.float32, device="cuda")
    D.dev_knn(x, y, z, k, m=m, nbr=nbr, rho=rho, rho_mode=rho_mode)
    D.dev_status()
    idx, d2 = O.knn(pos.astype(np.float64), k)
    got = nbr.cpu().numpy()
    assert np.array_equal(got[picks], idx[picks])
    assert np.array_equal(got, idx)
    want = O.sph_density(idx, d2, mass if rho_mode == 2 else None)
    assert np.array_equal(rho.cpu().numpy().view(np.int32), want.view(np.int32))

