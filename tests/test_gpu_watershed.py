"""GPU parity of the watershed of subhalo_den (galaxyfinder_amd/csrc/
watershed.hip): FindCoreDensity (subhaloden.mod6.c:1167-1475), the shell
thresholds (:2995-3024) and GetShellParticlesPeaks +
SaveRemainingParticles2LastShell (:572-678).

  * the reference's own outputs (fixture tests/golden/watershed_small.npz,
    written by the reference build; four cases, among them cores dropped by
    a final sweep that met another peak, whose particles then stay visited)
    through the host entry points and the device ones;
  * the oracle (pinned to the reference by tests/test_oracle_watershed.py)
    on the 131072-particle halo's own stage outputs (W4 kNN densities and
    32-NN lists, finddenpeak cores), on the 4194304-particle halo (core
    densities, members and the shells of the upper thresholds), and on two
    separate 280000-particle clumps with one core each, whose sweeps outgrow
    the LDS tier and the workgroup tier.
Integer and FP32 outputs formed by the reference's operations: the bar is
equality.
"""
import os
import sys
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from galaxyfinder_amd import density as D  # noqa: E402
from galaxyfinder_amd import halos  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden")
CASES = ["fdp", "rand", "grid", "ties"]


def say(msg):
    print("[case] " + msg, flush=True)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.fixture(scope="module")
def ws():
    f = np.load(os.path.join(GOLD, "watershed_small.npz"))
    pos = f["pos"].astype(np.float64)
    idx, _ = O.knn(pos, int(f["k"]))
    assert zlib.crc32(np.ascontiguousarray(idx, dtype=np.int64).tobytes()) == int(f["knn_crc"])
    return f, pos, idx


def _want_shells(f, name):
    pp, part, cp, cores = (f[name + "_" + s] for s in ("part_ptr", "part", "core_ptr", "cores"))
    return [(part[pp[i]:pp[i + 1]], cores[cp[i]:cp[i + 1]]) for i in range(len(pp) - 1)]


def _same_shells(got, want):
    assert len(got) == len(want)
    for i, ((a, b), (c, d)) in enumerate(zip(got, want)):
        assert np.array_equal(a, c), "shell %d particles" % i
        assert np.array_equal(b, d), "shell %d cores" % i


def _bits(a):
    return np.ascontiguousarray(a).view(np.int32)


@pytest.mark.parametrize("name", CASES)
def test_watershed_host_api_vs_reference_fixture(ws, name):
    f, pos, idx = ws
    den, peak = f[name + "_den"], f[name + "_peak"]
    n, nc = len(den), len(peak)
    say("host watershed fixture %s: %d cores" % (name, nc))
    bp = D.make_particles(pos, f["mass"], f["types"])
    core = np.zeros(nc, dtype=D.CORE_DTYPE)
    core["peak"] = peak
    haloid = np.zeros(n, np.int32)
    flag = np.zeros(n, np.uint8)
    nbr = np.ascontiguousarray(idx, dtype=np.int64)
    assert D.FindCoreDensity(bp, n, nbr, 32, core, nc, den, haloid, flag) == nc
    for key in ("coredensity", "nummem", "numstar", "starmass"):
        assert np.array_equal(_bits(core[key]), _bits(f[name + "_" + key])), key
    assert np.array_equal(haloid, f[name + "_haloid"])
    assert np.array_equal(flag, f[name + "_flag"])
    thr = D.shell_thresholds(den, n, core, nc)
    assert np.array_equal(_bits(thr), _bits(f[name + "_thr"]))
    shells = D.GetShellParticlesPeaks(n, nbr, 32, core, nc, den, haloid, flag, thr)
    _same_shells(shells, _want_shells(f, name))
    assert np.array_equal(flag, f[name + "_flag2"])
    assert np.array_equal(haloid, f[name + "_haloid2"])


def _dev_watershed(den, nbr, peak, types, mass, thr=None):
    """Device FindCoreDensity, thresholds and shells; numpy results."""
    import torch
    o = D.dev_find_core_density(den, nbr, peak, types, mass)
    D.dev_status()
    out = {k: v.cpu().numpy() for k, v in o.items()}
    if thr is None:
        thr = D.dev_shell_thresholds(den, o["coredensity"])
    flag, haloid = o["flag"].clone(), o["haloid"].clone()
    ns, pp, part, cp, cores = D.dev_shell_particles_peaks(den, nbr, peak, o["coredensity"], thr,
                                                          flag, haloid)
    torch.cuda.synchronize()
    pp, part, cp, cores = (t.cpu().numpy() for t in (pp, part, cp, cores))
    shells = [(part[pp[i]:pp[i + 1]], cores[cp[i]:cp[i + 1]]) for i in range(ns)]
    return out, thr, shells, flag.cpu().numpy(), haloid.cpu().numpy()


@pytest.mark.parametrize("name", CASES)
def test_watershed_device_api_vs_reference_fixture(ws, name):
    f, pos, idx = ws
    say("device watershed fixture %s" % name)
    o, thr, shells, fl, hid = _dev_watershed(_dev(f[name + "_den"]), _dev(idx.astype(np.int64)),
                                             _dev(f[name + "_peak"].astype(np.int32)),
                                             _dev(f["types"]), _dev(f["mass"]))
    for key in ("coredensity", "nummem", "numstar", "starmass"):
        assert np.array_equal(_bits(o[key]), _bits(f[name + "_" + key])), key
    assert np.array_equal(o["haloid"], f[name + "_haloid"])
    assert np.array_equal(o["flag"], f[name + "_flag"])
    assert np.array_equal(_bits(thr), _bits(f[name + "_thr"]))
    _same_shells(shells, _want_shells(f, name))
    assert np.array_equal(fl, f[name + "_flag2"]) and np.array_equal(hid, f[name + "_haloid2"])


def _stage(pos, mass, spec):
    import torch
    n, ng, k = len(pos), spec["nx"], 32
    g = D.grid_struct(ng, ng, ng, spec["xmin"], spec["ymin"], spec["zmin"], spec["cellsize"])
    x, y, z = (_dev(pos[:, a].astype(np.float32)) for a in range(3))
    den = torch.empty(D.padded_shape(ng, ng, ng), dtype=torch.float32, device="cuda")
    rt = torch.empty(n, dtype=torch.float32, device="cuda")
    nbr = torch.empty((n, k), dtype=torch.int64, device="cuda")
    rk = torch.empty(n, dtype=torch.float32, device="cuda")
    D.dev_density_stage(x, y, z, _dev(mass.astype(np.float32)), g, spec["RG"], k, den, rt, nbr,
                        rk)
    D.dev_status()
    return nbr, rk


def _check_vs_oracle(den_d, nbr_d, peak, types, mass, thr_count=None, tag=""):
    den, nbr = den_d.cpu().numpy(), nbr_d.cpu().numpy()
    o = O.find_core_density(den, nbr, peak, types, mass)
    thr = O.shell_thresholds(den, o["coredensity"])
    if thr_count is not None:
        thr = thr[:thr_count]
    g, gthr, gshells, gfl, ghid = _dev_watershed(den_d, nbr_d, _dev(peak), _dev(types),
                                                 _dev(mass), thr=thr)
    for key in ("coredensity", "nummem", "numstar", "starmass"):
        assert np.array_equal(_bits(g[key]), _bits(o[key])), key
    assert np.array_equal(g["haloid"], o["haloid"]) and np.array_equal(g["flag"], o["flag"])
    shells, fl, hid = O.core_shells(den, nbr, peak, o["coredensity"], thr, o["flag"], o["haloid"])
    say("%s: %d cores, members %d (max %d), %d shells" % (
        tag, len(peak), int(o["nummem"].sum()), int(o["nummem"].max()), len(shells)))
    _same_shells(gshells, shells)
    assert np.array_equal(gfl, fl) and np.array_equal(ghid, hid)


def test_watershed_on_the_131072_particle_halo():
    pos, mass, spec = halos.single_halo_config(1)
    nbr, rk = _stage(pos, mass, spec)
    n = len(pos)
    types = np.where(np.arange(n) % 4 == 0, 1, 2).astype(np.int32)
    rk_h, nbr_h = rk.cpu().numpy(), nbr.cpu().numpy()
    peak = O.finddenpeak(rk_h, nbr_h, types, 0, float(np.quantile(rk_h, 0.5))).astype(np.int32)
    _check_vs_oracle(rk, nbr, peak, types, mass.astype(np.float64), tag="C1 finddenpeak cores")


def test_watershed_on_the_4194304_particle_halo():
    """Core densities and members of the 38k finddenpeak cores; the shells
    of the eight upper thresholds (the lower ones sweep millions of
    particles per root, which the oracle needs minutes for; the two-clump
    test covers such sweeps)."""
    pos, mass, spec = halos.single_halo_config(2)
    nbr, rk = _stage(pos, mass, spec)
    n = len(pos)
    types = np.where(np.arange(n) % 4 == 0, 1, 2).astype(np.int32)
    rk_h, nbr_h = rk.cpu().numpy(), nbr.cpu().numpy()
    peak = O.finddenpeak(rk_h, nbr_h, types, 0, float(np.quantile(rk_h, 0.5))).astype(np.int32)
    _check_vs_oracle(rk, nbr, peak, types, mass.astype(np.float64), thr_count=8,
                     tag="C2 finddenpeak cores")


def _clump(rng, n, center, a):
    # Plummer sphere truncated at 5a
    r = a / np.sqrt(rng.uniform(0.02, 1.0, n) ** (-2.0 / 3.0) - 1.0)
    r = np.minimum(r, 5 * a)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return center + r[:, None] * v


def test_watershed_sweeps_beyond_the_lds_and_workgroup_tiers():
    """Two clumps far apart, one core each at its densest particle: no sweep
    reaches the other core, so every threshold trial passes and the final
    sweeps hold whole clumps (280000 particles: past the LDS tier of 448 and
    the workgroup tier of 64512); the shells' lower thresholds sweep them
    too.  Plus 30 extra cores at random dense particles of the first clump."""
    rng = np.random.default_rng(77)
    m = 280000
    pos = np.concatenate([_clump(rng, m, np.array([0.3, 0.3, 0.3]), 0.03),
                          _clump(rng, m, np.array([0.7, 0.7, 0.7]), 0.03)]).astype(np.float32)
    pos = pos.astype(np.float64)
    n = len(pos)
    spec = dict(nx=64, ny=64, nz=64, xmin=0.0, ymin=0.0, zmin=0.0, cellsize=1.0 / 64,
                RG=2.0 / 64)
    mass = np.ones(n, np.float32)
    nbr, rk = _stage(pos, mass, spec)
    rk_h = rk.cpu().numpy()
    types = np.where(rng.random(n) < 0.5, 1, 2).astype(np.int32)
    mass64 = rng.uniform(0.5, 1.5, n)
    peak = np.array([int(np.argmax(rk_h[:m])), m + int(np.argmax(rk_h[m:]))], np.int32)
    _check_vs_oracle(rk, nbr, peak, types, mass64, tag="two clumps, two cores")
    med = float(np.median(rk_h[:m]))
    extra = rng.choice(np.nonzero(rk_h[:m] > med)[0], 30, replace=False).astype(np.int32)
    _check_vs_oracle(rk, nbr, np.concatenate([peak, extra]), types, mass64,
                     tag="two clumps, 32 cores")
