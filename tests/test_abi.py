"""CPU checks of the C-ABI boundary and the host-side logic (no GPU calls):
the library loads, exports every function include/pgalf_density.h declares,
the particle record matches the reference's 88-byte layout, and the Python
wrappers reject malformed arguments before touching the device."""
import ctypes
import os
import re
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from galaxyfinder_amd import _lib, density as D, halos  # noqa: E402

HEADERS = [os.path.join(ROOT, "include", h) for h in ("pgalf_density.h", "pgalf_catalog.h")]


def declared_functions():
    text = "".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "sizeof", "defined"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert "assign_density_TSC" in names and "gaussian_Smoothing" in names
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), "missing export %s" % n
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of date"


def test_version_and_error_state_without_gpu_calls():
    lib = _lib.load()
    assert b"gfx950" in lib.pgalf_version()
    assert lib.pgalf_last_error() == 0


def test_particle_record_layout():
    assert D.PARTICLE_DTYPE.itemsize == 88
    off = {k: v[1] for k, v in D.PARTICLE_DTYPE.fields.items()}
    assert off["type"] == 0 and off["mass"] == 8 and off["x"] == 16
    assert off["y"] == 24 and off["z"] == 32 and off["indx"] == 72 and off["bp"] == 80
    bp = D.make_particles(np.array([[0.1, 0.2, 0.3]]), np.array([2.0]), np.array([2]))
    raw = bp.tobytes()
    assert np.frombuffer(raw[16:40], dtype=np.float64).tolist() == [0.1, 0.2, 0.3]
    assert np.frombuffer(raw[8:16], dtype=np.float64)[0] == 2.0


def test_padded_shape():
    assert D.padded_shape(37, 35, 32) == (32, 35, 38)
    assert D.padded_shape(64, 64, 64) == (64, 64, 66)


def test_argument_checks_raise_before_device():
    bp = D.make_particles(np.zeros((4, 3)))
    with pytest.raises(TypeError):
        D.assign_density_TSC(bp, 4, np.zeros(10, dtype=np.float64), 4, 4, 4, 0, 0, 0, 1.0, 5)
    with pytest.raises(ValueError):
        D.findDen(bp, 4, np.zeros(4, dtype=np.float32), 0, 0, 0, 1.0,
                  np.zeros(5, dtype=np.float32), 4, 4, 4)
    with pytest.raises(TypeError):
        D.findsphdensity(bp, 4, np.zeros(4 * 3, dtype=np.int32), 3,
                         np.zeros(4, dtype=np.float32))


def test_halo_generator_seeded_and_inside_box():
    a, ma = halos.small_halo(5000, seed=1)
    b, mb = halos.small_halo(5000, seed=1)
    assert np.array_equal(a, b) and np.array_equal(ma, mb)
    assert a.dtype == np.float32 and ma.dtype == np.float32
    assert a.min() >= 0.1 and a.max() <= 0.9
    p, m, spec = halos.single_halo_config(1)
    assert len(p) == 131072 and spec["nx"] == 64 and spec["k"] == 32
    assert p.min() >= 0.1 - 1e-6 and p.max() <= 0.9 + 1e-6
    assert abs(m.astype(np.float64).sum() - 1.0) < 1e-4


def test_batch_create_validates_before_device():
    """pgalf_batch_create checks the catalogue on the host: a halo with fewer
    than k particles is a short list (the reference exits with 99 there),
    a malformed offset table a bad argument."""
    grids = [(32, 32, 32, 0.0, 0.0, 0.0, 1.0 / 32)] * 2
    with pytest.raises(D.PgalfError) as ei:
        D.Batch(np.array([0, 100, 120], dtype=np.int64), grids, [2.0 / 32] * 2, 32)
    assert ei.value.code == 5
    with pytest.raises(D.PgalfError) as ei:
        D.Batch(np.array([5, 100, 220], dtype=np.int64), grids, [2.0 / 32] * 2, 32)
    assert ei.value.code == 1
    with pytest.raises(D.PgalfError) as ei:
        D.Batch(np.array([0, 100, 220], dtype=np.int64), grids, [2.0 / 32] * 2, 33)
    assert ei.value.code == 1


def test_catalogue_generator():
    p, m, off, grids, rg = halos.catalogue(config_id=103, nhalo=50, nmax=20000)
    c = np.diff(off)
    assert off[0] == 0 and off[-1] == len(p) == len(m) and len(grids) == len(rg) == 50
    assert c.min() >= 100 and c.max() <= 20000
    assert p.min() >= 0.1 - 1e-6 and p.max() <= 0.9 + 1e-6
    for n, g, r in zip(c, grids, rg):
        assert g[0] == halos.grid_for(n) and abs(r * g[0] - 2.0) < 1e-6
    assert halos.grid_for(100) == 32 and halos.grid_for(1048576) == 256
    assert halos.grid_for(32768) == 128
    # dN/dn ~ n^-2 on [100, 1e6]: the median is ~2 nmin
    big = halos.powerlaw_counts(halos.rng_for(0), 20000, 2.0, 100, 1048576)
    assert 180 < np.median(big) < 220


def test_vectorised_catalogue():
    p, m, off, grids, rg = halos.catalogue_vec(config_id=105, nhalo=300, nmax=50000)
    c = np.diff(off)
    assert off[0] == 0 and off[-1] == len(p) == len(m) and len(grids) == len(rg) == 300
    assert c.min() >= 64 and c.max() <= 50000
    assert p.min() >= 0.1 - 1e-6 and p.max() <= 0.9 + 1e-6
    for h in (0, int(np.argmax(c))):
        assert np.allclose(m[off[h]:off[h + 1]], 1.0 / c[h])
        # each halo is centred in its own box
        assert np.abs(p[off[h]:off[h + 1]].mean(axis=0) - 0.5).max() < 0.15


@pytest.mark.skipif(not os.path.isdir("/root/reference/NewGalFinder"),
                    reason="needs the reference sources (build container only)")
def test_reference_nnost_links_against_the_library():
    """the reference's own nnost.o, linked without
    tsc_omp2.o and gsmooth.o, takes assign_density_TSC, gaussian_Smoothing
    and findDen from libpgalf_density.so (link only, no call)."""
    import subprocess
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-B", "linkcheck"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for sym in ("assign_density_TSC", "gaussian_Smoothing", "findDen"):
        assert "nnost.o: reference to %s" % sym in r.stderr
        assert "libpgalf_density.so: definition of %s" % sym in r.stderr


def test_merging_peak_host_checks_without_gpu_calls():
    """pgalf_merging_peak returns 0 for no cores and -1 (PGALF_ERR_ARG) for
    a peak outside the particles, before any device work."""
    bp = D.make_particles(np.array([[0.1, 0.2, 0.3], [0.4, 0.5, 0.6]]))
    core = np.zeros(2, dtype=D.CORE_DTYPE)
    assert D.MergingPeak(bp, 2, core, 0) == 0
    core["peak"] = [0, 7]
    with pytest.raises(_lib.PgalfError) as e:
        D.MergingPeak(bp, 2, core, 2)
    assert e.value.code == 1
    with pytest.raises(_lib.PgalfError):
        D.MergingPeak(bp, 2, core, 2, iflag=3)


def test_catalog_shell_layout():
    """pgalf_catalog_shell_layout: halo h's pointer entries (cores x
    nshelldivide + 2, nshelldivide 10 up to 1000 particles, else
    min(15, 10 log10(1.5 n))) and core-entry room (16 x cores x (nsd + 1) +
    16), laid out halo after halo."""
    import ctypes
    lib = _lib.load()
    off = np.array([0, 500, 1500, 101500], dtype=np.int64)
    coff = np.array([0, 3, 3, 10], dtype=np.int64)
    sb, cb, cc = (np.zeros(3, np.int64) for _ in range(3))
    tot = ctypes.c_longlong(0)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    ns = lib.pgalf_catalog_shell_layout(3, p(off), p(coff), p(sb), p(cb), p(cc),
                                        ctypes.byref(tot))
    nsd = [10, int(10 * np.log10(1000 * 1.5)), 15]
    want_s = [3 * nsd[0] + 2, 0 * nsd[1] + 2, 7 * nsd[2] + 2]
    want_c = [16 * c * (d + 1) + 16 for c, d in zip((3, 0, 7), nsd)]
    assert list(sb) == [0, want_s[0], want_s[0] + want_s[1]] and ns == sum(want_s)
    assert list(cc) == want_c and list(cb) == [0, want_c[0], want_c[0] + want_c[1]]
    assert tot.value == sum(want_c)
