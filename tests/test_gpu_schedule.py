"""GPU checks of the launch schedule (pgalf_set_schedule):

  * the bin knobs (packed kNN query on/off, XCD chunk, deposit unit size)
    leave every output bit unchanged: the single-halo stage (deposit,
    smoothed grid, read-back, 32-NN lists, W4 densities) and a batch created
    under a non-default schedule equal the defaults' byte for byte;
  * smooth_direct_max moves lines of 40, 48 and 56 cells between the direct
    convolution and the FFT: the smoothed grids agree within 1e-5 normwise
    (both methods evaluate the reference's Gaussian filter, gsmooth.c:76-86,
    to FP32 rounding), the read-back densities likewise;
  * the tuner runs end to end on a small bin and returns a valid document.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from galaxyfinder_amd import density as D, halos, tune  # noqa: E402

pytestmark = pytest.mark.gpu
K = 32


def say(msg):
    print("[case] " + msg, flush=True)


@pytest.fixture(autouse=True)
def _restore():
    yield
    D.set_schedule()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _stage(pos, mass, ng):
    import torch
    n = len(pos)
    g = D.grid_struct(ng, ng, ng, 0.0, 0.0, 0.0, 1.0 / ng)
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    den = torch.empty(D.padded_shape(ng, ng, ng), dtype=torch.float32, device="cuda")
    rt = torch.empty(n, dtype=torch.float32, device="cuda")
    rk = torch.empty(n, dtype=torch.float32, device="cuda")
    nbr = torch.empty((n, K), dtype=torch.int64, device="cuda")
    D.dev_density_stage(x, y, z, _dev(mass), g, 2.0 / ng, K, den, rt, nbr, rk)
    D.dev_status()
    return [t.cpu().numpy() for t in (den, rt, nbr, rk)]


SCHEDULES = [
    dict(knn_packed=0, knn_xcd_chunk=0, tsc_unit_entries=1024),
    dict(knn_packed=1, knn_xcd_chunk=16, tsc_unit_entries=256),
    dict(knn_packed=1, knn_xcd_chunk=1024, tsc_unit_entries=65536),
]


@pytest.mark.parametrize("n,ng", [(30000, 64), (200000, 64)])
def test_bin_knobs_leave_the_stage_bit_identical(n, ng):
    pos, mass = halos.small_halo(n, seed=7)
    ref = _stage(pos, mass, ng)
    for sc in SCHEDULES:
        say("stage n=%d under %s" % (n, sc))
        D.set_schedule([dict(nmax=n, **sc)])
        got = _stage(pos, mass, ng)
        for a, b in zip(got, ref):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _batch(off, grids, rg, x, y, z, m):
    import torch
    n = int(off[-1])
    b = D.Batch(off, grids, rg, K)
    den = torch.empty(b.grid_floats, dtype=torch.float32, device="cuda")
    rt = torch.empty(n, dtype=torch.float32, device="cuda")
    rk = torch.empty(n, dtype=torch.float32, device="cuda")
    nbr = torch.empty((n, K), dtype=torch.int64, device="cuda")
    b.run(x, y, z, m, den, rt, nbr, rk)
    b.status()
    offs = b.grid_offsets
    b.close()
    return [t.cpu().numpy() for t in (den, rt, nbr, rk)], offs


def test_batch_keeps_the_schedule_it_was_created_under():
    pos, mass, off, grids, rg = halos.catalogue_vec(5, nhalo=600, nmin=64, nmax=65536)
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(mass)
    ref, _ = _batch(off, grids, rg, x, y, z, m)
    for sc in SCHEDULES:
        say("batch of 600 halos under %s" % sc)
        D.set_schedule([dict(nmax=1 << 40, **sc)])
        got, _ = _batch(off, grids, rg, x, y, z, m)
        for a, b in zip(got, ref):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_smoothing_threshold_within_tolerance():
    nhalo = 100
    pos, mass, off, _, _ = halos.catalogue_vec(5, nhalo=nhalo, nmin=64, nmax=8192)
    sides = [tune.GRID_SIDES[h % len(tune.GRID_SIDES)] for h in range(nhalo)]
    grids = [(g, g, g, 0.0, 0.0, 0.0, 1.0 / g) for g in sides]
    rg = np.array([2.0 / g for g in sides], dtype=np.float32)
    x, y, z = (_dev(pos[:, a]) for a in range(3))
    m = _dev(mass)
    ref, offs = _batch(off, grids, rg, x, y, z, m)
    for dm in (32, 64):
        say("batch of %d halos, smooth_direct_max=%d" % (nhalo, dm))
        D.set_schedule([], dm)
        got, _ = _batch(off, grids, rg, x, y, z, m)
        assert np.array_equal(got[2], ref[2])  # lists do not depend on the grid
        worst = 0.0
        for h in range(nhalo):
            a, e = offs[h], offs[h + 1]
            r = ref[0][a:e].astype(np.float64)
            worst = max(worst, np.linalg.norm(got[0][a:e] - r) / np.linalg.norm(r))
            p0, p1 = off[h], off[h + 1]
            rr = ref[1][p0:p1].astype(np.float64)
            worst = max(worst, np.linalg.norm(got[1][p0:p1] - rr) / np.linalg.norm(rr))
        say("  worst normwise difference %.2e" % worst)
        assert worst <= 1e-5


def test_tuner_runs_on_a_small_bin():
    doc = tune.tune(bins=((16384, 8192),), reps=1, smooth_halos=50, log=say)
    bins, dm = tune.validate(doc)
    assert len(bins) == 1 and bins[0]["nmax"] == 16384 and dm in tune.DIRECT_MAX
    tune.apply(doc)
    assert D.get_schedule(8192)[0] == bins[0]
