"""GPU parity of the catalogue watershed (pgalf_dev_catalog_watershed):
FindCoreDensity, the shell thresholds and the shells of every halo of a
catalogue in one call, against

  * the one-halo device entry points (pgalf_dev_find_core_density,
    pgalf_dev_shell_thresholds, pgalf_dev_shell_particles_peaks, themselves
    bit-exact against the reference's FindCoreDensity and
    GetShellParticlesPeaks in tests/test_gpu_watershed.py) called halo by
    halo on the same arrays -- every output of every halo equal;
  * the oracle (pinned to the reference by tests/test_oracle_watershed.py)
    on a sample of halos.

The catalogue: 160 power-law halos (100 to 20,000 particles) through the
batched stage (halo-local 32-NN lists, W4 densities), finddenpeak cores above
each halo's 30% density quantile (plus, in every 8th halo, three dense
particles that are not local maxima, so some cores are dropped), FP64
masses, one particle in three a star.  The
workgroup path takes the halos up to small_max, the one-halo path the rest
on four host threads at once: small_max 3,000 (both paths), the default
(every halo on the workgroup path) and 1 (every halo on the one-halo path,
160 halos over the four threads).  Integer and FP32 outputs: the bar is
equality.
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


def say(msg):
    print("[case] " + msg, flush=True)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.fixture(scope="module")
def cat():
    import torch
    pos, mass, off, grids, rg = halos.catalogue(config_id=11, nhalo=160, nmin=100, nmax=20000)
    n = int(off[-1])
    say("catalogue: %d halos, %d particles, largest %d" % (len(off) - 1, n, np.diff(off).max()))
    b = D.Batch(off, grids, rg, 32)
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(mass)
    den = torch.empty(b.grid_floats, dtype=torch.float32, device="cuda")
    nbr = torch.empty((n, 32), dtype=torch.int64, device="cuda")
    rk = torch.empty(n, dtype=torch.float32, device="cuda")
    b.run(x, y, z, m, den, None, nbr, rk)
    b.status()
    rk_h, nbr_h = rk.cpu().numpy(), nbr.cpu().numpy()
    rng = np.random.default_rng(5)
    types = np.where(rng.random(n) < 1.0 / 3.0, 1, 2).astype(np.int32)
    mass64 = rng.uniform(0.5, 2.0, n) * 1.0e6
    peaks, coff = [], [0]
    for h in range(len(off) - 1):
        a, e = off[h], off[h + 1]
        d = rk_h[a:e]
        pk = O.finddenpeak(d, nbr_h[a:e], types[a:e], 0, float(np.quantile(d, 0.3)))
        if h % 8 == 0:
            # three dense particles that are not local maxima: their final
            # sweeps can meet another peak (dropped cores, visited marks)
            rest = np.setdiff1d(np.nonzero(d > np.median(d))[0], pk)
            pk = np.concatenate([pk, rng.choice(rest, min(3, len(rest)), replace=False)])
        peaks.append(pk.astype(np.int32))
        coff.append(coff[-1] + len(pk))
    peak = np.concatenate(peaks)
    return dict(off=off, coff=np.array(coff, np.int64), rk=rk, nbr=nbr, rk_h=rk_h,
                nbr_h=nbr_h, types=types, mass64=mass64, peak=peak)


def _run_catalog(c, small_max):
    o = D.dev_catalog_watershed(c["off"], c["coff"], c["rk"], c["nbr"], _dev(c["peak"]),
                                _dev(c["types"]), _dev(c["mass64"]), small_max=small_max)
    return {k_: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k_, v in o.items()}


def _one_halo(c, h):
    """The one-halo device path on halo h's slices."""
    import torch
    a, e = int(c["off"][h]), int(c["off"][h + 1])
    ca, ce = int(c["coff"][h]), int(c["coff"][h + 1])
    den = c["rk"][a:e].contiguous()
    nbr = c["nbr"][a:e].contiguous()
    peak = _dev(c["peak"][ca:ce])
    o = D.dev_find_core_density(den, nbr, peak, _dev(c["types"][a:e]), _dev(c["mass64"][a:e]))
    thr = D.dev_shell_thresholds(den, o["coredensity"])
    fl, hid = o["flag"].clone(), o["haloid"].clone()
    ns, pp, part, cp, cores = D.dev_shell_particles_peaks(den, nbr, peak, o["coredensity"], thr,
                                                          fl, hid)
    torch.cuda.synchronize()
    pp, part, cp, cores = (t.cpu().numpy() for t in (pp, part, cp, cores))
    shells = [(part[pp[i]:pp[i + 1]], cores[cp[i]:cp[i + 1]]) for i in range(ns)]
    core = {k_: o[k_].cpu().numpy() for k_ in ("coredensity", "nummem", "numstar", "starmass")}
    return core, fl.cpu().numpy(), hid.cpu().numpy(), shells


def _same_halo(got, c, h, core, fl, hid, shells):
    a, e = int(c["off"][h]), int(c["off"][h + 1])
    ca, ce = int(c["coff"][h]), int(c["coff"][h + 1])
    for key in ("coredensity", "nummem", "numstar", "starmass"):
        assert np.array_equal(np.ascontiguousarray(got[key][ca:ce]).view(np.int32),
                              np.ascontiguousarray(core[key]).view(np.int32)), (h, key)
    assert np.array_equal(got["flag"][a:e], fl), h
    assert np.array_equal(got["haloid"][a:e], hid), h
    gs = D.catalog_shells_of(got, c["off"], h)
    assert len(gs) == len(shells), h
    for (p, q), (r, s) in zip(gs, shells):
        assert np.array_equal(p, r) and np.array_equal(q, s), h


@pytest.mark.parametrize("small_max", [3000, 0, 1])
def test_catalog_watershed_equals_the_one_halo_path(cat, small_max):
    got = _run_catalog(cat, small_max)
    nh = len(cat["off"]) - 1
    assert not got["status"].any()
    say("catalogue watershed small_max=%d: %d cores (%d dropped), %d shells" % (
        small_max, int(cat["coff"][-1]), int(np.sum(got["nummem"][:int(cat["coff"][-1])] == 0)),
        int(got["nshell"].sum())))
    for h in range(nh):
        core, fl, hid, shells = _one_halo(cat, h)
        _same_halo(got, cat, h, core, fl, hid, shells)


def test_catalog_watershed_against_the_oracle(cat):
    got = _run_catalog(cat, 3000)
    sizes = np.diff(cat["off"])
    pick = list(np.argsort(-sizes)[:3]) + list(range(0, len(sizes), 40))
    for h in pick:
        a, e = int(cat["off"][h]), int(cat["off"][h + 1])
        ca, ce = int(cat["coff"][h]), int(cat["coff"][h + 1])
        den, nbr = cat["rk_h"][a:e], cat["nbr_h"][a:e]
        peak = cat["peak"][ca:ce]
        o = O.find_core_density(den, nbr, peak, cat["types"][a:e], cat["mass64"][a:e])
        thr = O.shell_thresholds(den, o["coredensity"])
        shells, fl, hid = O.core_shells(den, nbr, peak, o["coredensity"], thr, o["flag"],
                                        o["haloid"])
        say("halo %d: %d particles, %d cores, %d shells" % (h, e - a, ce - ca, len(shells)))
        _same_halo(got, cat, h, o, fl, hid, shells)


def test_catalog_watershed_on_the_10000_halo_catalogue():
    """The 10,000-halo catalogue (9.15M particles, 100 to 451,434 per halo)
    through the batched stage; each halo's finddenpeak cores above its
    median density (42k cores).  One call at the default split (halos over
    131,072 particles through the one-halo path on four host threads, the
    rest on the workgroup path); against the one-halo entry points on the
    eight largest halos, the eight around the split and 24 seeded ones."""
    import torch
    pos, mass, off, grids, rg = halos.catalogue(config_id=3)
    n, nh = int(off[-1]), len(off) - 1
    b = D.Batch(off, grids, rg, 32)
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(mass)
    den = torch.empty(b.grid_floats, dtype=torch.float32, device="cuda")
    nbr = torch.empty((n, 32), dtype=torch.int64, device="cuda")
    rk = torch.empty(n, dtype=torch.float32, device="cuda")
    b.run(x, y, z, m, den, None, nbr, rk)
    b.status()
    del den
    b.close()
    coff, peaks, types = halos.catalogue_cores(off, rk, nbr)
    mass64 = m.double()
    peak = np.concatenate(peaks)
    c = dict(off=off, coff=coff, rk=rk, nbr=nbr, types=types.cpu().numpy(),
             mass64=mass64.cpu().numpy(), peak=peak)
    got = _run_catalog(c, 0)
    assert not got["status"].any()
    sizes = np.diff(off)
    order = np.argsort(-sizes, kind="stable")
    split = int(np.searchsorted(-sizes[order], -131072, side="left"))
    pick = list(order[:8]) + list(order[max(0, split - 4):split + 4])
    pick += list(np.random.default_rng(3).choice(nh, 24, replace=False))
    say("10k catalogue: %d cores, %d shells; checking %d halos" % (
        int(coff[-1]), int(got["nshell"].sum()), len(pick)))
    for h in dict.fromkeys(int(v) for v in pick):
        core, fl, hid, shells = _one_halo(c, h)
        _same_halo(got, c, h, core, fl, hid, shells)


def test_catalog_watershed_with_halos_without_cores(cat):
    """Halos with no cores (the reference's subhalo_den never gets here with
    none; a caller of the catalogue call may pass them): no core outputs,
    one shell holding every particle in index order, the same as the
    one-halo path; the other halos as before.  Both paths."""
    import torch
    off, coff, peak = cat["off"], cat["coff"], cat["peak"]
    keep = np.ones(len(off) - 1, bool)
    keep[[1, 7, 50, 100]] = False
    peaks = [peak[coff[h]:coff[h + 1]] if keep[h] else peak[:0] for h in range(len(off) - 1)]
    c2 = dict(cat)
    c2["peak"] = np.concatenate(peaks).astype(np.int32)
    c2["coff"] = np.concatenate([[0], np.cumsum([len(p) for p in peaks])]).astype(np.int64)
    for small_max in (0, 1):
        got = _run_catalog(c2, small_max)
        assert not got["status"].any()
        for h in (0, 1, 7, 8, 50, 100):
            core, fl, hid, shells = _one_halo(c2, h)
            _same_halo(got, c2, h, core, fl, hid, shells)
            if not keep[h]:
                n = int(off[h + 1] - off[h])
                gs = D.catalog_shells_of(got, off, h)
                assert len(gs) == 1 and np.array_equal(gs[0][0], np.arange(n))
                assert len(gs[0][1]) == 0
    say("halos without cores: equal to the one-halo path")
