"""CPU tests of the FoF catalogue loader and the GALFIND writer
(galaxyfinder_amd/csrc/catalog.hip through include/pgalf_catalog.h).

  * record layouts: the numpy dtypes the tests write with and the library's
    fixed sizes equal the reference build's sizeof / offsetof (fixture
    tests/golden/catalog_small.npz, written from FREAD / Dump2BP /
    mklocalize / write_data of gfind.c and subhaloden.mod6.c compiled in
    place);
  * loader: localised positions, masses, types, link02 and Dump2BP origins
    bit-identical to the reference's for every halo, the two that span the
    periodic box included (PGALF_UNWRAP_REFERENCE); with
    PGALF_UNWRAP_SORTED those two come out compact and the others are
    unchanged; batches by halo count and particle budget cover the halos of
    at least 30 particles in order; each halo's grid is the one
    pgalf_lag_density_stage sizes (nnost.c:1040-1058; checked on the GPU in
    tests/test_gpu_catalog.py);
  * writer: GALFIND.DATA, GALCATALOG.LIST and the background file
    byte-identical to the reference's write_data.
"""
import hashlib
import os
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

import catalog_gen as C  # noqa: E402
from galaxyfinder_amd import catalog as CAT  # noqa: E402

GOLD = os.path.join(ROOT, "tests", "golden", "catalog_small.npz")


@pytest.fixture(scope="module")
def files():
    d = tempfile.mkdtemp()
    return C.write(d)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_record_layouts_match_the_reference(gold):
    lay = gold["layouts"].tolist()
    assert lay[:7] == [C.HALOQ.itemsize, C.PM.itemsize, C.GAS.itemsize, C.SINK.itemsize,
                       C.PM.itemsize, C.HALOINFO.itemsize, C.SUBINFO.itemsize]
    f = lambda dt, k: dt.fields[k][1]  # noqa: E731
    assert lay[7:] == [f(C.HALOQ, "x"), f(C.HALOQ, "mass"), f(C.PM, "mass"), f(C.PM, "id"),
                       f(C.PM, "tp"), f(C.PM, "partp"), f(C.GAS, "vx"), f(C.GAS, "temp"),
                       f(C.GAS, "mass"), f(C.GAS, "potent"), f(C.SINK, "mass"),
                       f(C.SINK, "id"), f(C.HALOINFO, "totm"), f(C.SUBINFO, "totm"),
                       f(C.HALOQ, "vx")]
    assert CAT.HALOQ_DTYPE.itemsize == 96


def _all(files, unwrap=CAT.UNWRAP_REFERENCE, **kw):
    cat = CAT.Catalog(*files, unwrap=unwrap)
    b = cat.next_batch(min_np=0, **kw)
    cat.close()
    return b


def test_loader_equals_reference_localisation(files, gold):
    b = _all(files)
    assert np.array_equal(b["offsets"], gold["offsets"])
    for k in ("x", "y", "z", "mass", "type", "link02"):
        assert np.array_equal(b[k], gold[k]), k
    assert np.array_equal(b["origin"], gold["origin"])
    # the reference's unwrap leaves the halos that span the box wide open
    ext = [np.ptp(b["x"][a:e]) for a, e in zip(b["offsets"][:-1], b["offsets"][1:])]
    assert ext[1] > 90.0 and max(ext[0], ext[3], ext[4]) < 5.0


def test_sorted_unwrap_makes_spanning_halos_compact(files):
    r = _all(files)
    s = _all(files, unwrap=CAT.UNWRAP_SORTED)
    off = r["offsets"]
    for h in range(len(off) - 1):
        a, e = off[h], off[h + 1]
        ext = max(np.ptp(s[k][a:e]) for k in "xyz")
        assert ext < 5.0, h
        if h in (0, 3, 4):  # inside the box: both modes agree
            for k in "xyz":
                assert np.array_equal(s[k][a:e], r[k][a:e])
    g = s["grids"]
    assert int(g["nx"].min()) > 10 and int(g["nx"].max()) < 1000


def test_batches_cover_the_halos_of_at_least_30_particles(files):
    full = _all(files)
    sizes = np.diff(full["offsets"])
    keep = [h for h in range(len(sizes)) if sizes[h] >= CAT.MIN_HALO_NP]
    cat = CAT.Catalog(*files)
    got = []
    while True:
        b = cat.next_batch(max_halos=2, max_particles=900)
        if b is None:
            break
        nb = len(b["offsets"]) - 1
        assert nb <= 2 and (nb == 1 or b["offsets"][-1] <= 900)
        got.append(b)
    cat.close()
    halos = []
    for b in got:
        for i in range(len(b["offsets"]) - 1):
            halos.append((b, i))
    assert len(halos) == len(keep)
    for (b, i), h in zip(halos, keep):
        a, e = b["offsets"][i], b["offsets"][i + 1]
        fa, fe = full["offsets"][h], full["offsets"][h + 1]
        for k in ("x", "y", "z", "mass", "type"):
            assert np.array_equal(b[k][a:e], full[k][fa:fe])
        assert b["haloq"][i]["np"] == sizes[h]
        assert b["grids"][i] == full["grids"][h]
        ra, re_ = b["record_offsets"][i], b["record_offsets"][i + 1]
        fra, fre = full["record_offsets"][h], full["record_offsets"][h + 1]
        assert np.array_equal(b["records"][ra:re_], full["records"][fra:fre])


def test_writer_equals_reference_write_data(files, gold):
    b = _all(files)
    off, roff = b["offsets"], b["record_offsets"]
    d = tempfile.mkdtemp()
    outs = [os.path.join(d, f) for f in ("GALFIND.DATA", "GALCATALOG.LIST", "background")]
    w = CAT.Writer(*outs)
    for h in range(len(off) - 1):
        a, e = off[h], off[h + 1]
        if e - a < CAT.MIN_HALO_NP:
            continue
        mpeak, p2 = C.membership(int(e - a))
        assert mpeak == gold["mpeak"][h]
        w.halo(b["type"][a:e], b["records"][roff[h]:roff[h + 1]], p2, mpeak)
    w.close()
    data = [open(f, "rb").read() for f in outs]
    assert [len(x) for x in data] == gold["nbytes"].tolist()
    assert [hashlib.sha256(x).hexdigest() for x in data] == gold["sha256"].tolist()
    lst = np.frombuffer(data[1], dtype=np.uint8)
    assert np.array_equal(lst, gold["catalog_list"])
    # GALCATALOG.LIST: HaloInfo then its SubInfos; subhalo counts add up
    info = np.frombuffer(data[1][:112], dtype=C.HALOINFO)[0]
    assert info["nsub"] == gold["mpeak"][0] and info["npall"] == off[1] - off[0]


def test_writer_and_loader_argument_errors():
    from galaxyfinder_amd._lib import PgalfError
    with pytest.raises(PgalfError):
        CAT.Catalog("/nonexistent/a", "/nonexistent/b")
    d = tempfile.mkdtemp()
    w = CAT.Writer(os.path.join(d, "a"), os.path.join(d, "b"))
    with pytest.raises(PgalfError):
        w.halo(np.array([7], dtype=np.int32), np.zeros(136, dtype=np.uint8),
               np.zeros(1, dtype=np.int32), 1)
    w.close()


def test_reference_build_live(files, gold):
    """The same comparison against the reference build itself, when it is
    present (it is in the build container)."""
    from oracle import oracle as O
    import ctypes
    R = O.ref_lib()
    if R is None:
        pytest.skip("reference build oracle/_ref absent")
    R.ref_init(4 << 30)
    b = _all(files)
    n = int(b["offsets"][-1])
    x, y, z, m = (np.zeros(n) for _ in range(4))
    ty = np.zeros(n, dtype=np.int32)
    l2 = np.zeros(n, dtype=np.float32)
    off = np.zeros(64, dtype=np.int64)
    org = np.zeros(192)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    R.ref_read_catalog.restype = ctypes.c_longlong
    nh = R.ref_read_catalog(files[0].encode(), files[1].encode(), 63, ctypes.c_longlong(n),
                            p(x), p(y), p(z), p(m), p(ty), p(l2), p(off), p(org))
    assert nh == len(b["offsets"]) - 1
    assert np.array_equal(x, b["x"]) and np.array_equal(l2, b["link02"])


def test_one_batch_skipping_small_halos(files):
    """The halos of at least 30 particles in one batch: the same particles,
    records and grids as the unfiltered read minus the small halo (the
    member records of a skipped halo lie between those of kept ones)."""
    full = _all(files)
    cat = CAT.Catalog(*files)
    b = cat.next_batch()
    cat.close()
    sizes = np.diff(full["offsets"])
    keep = [h for h in range(len(sizes)) if sizes[h] >= CAT.MIN_HALO_NP]
    assert len(keep) < len(sizes) and len(b["offsets"]) - 1 == len(keep)
    for i, h in enumerate(keep):
        a, e = b["offsets"][i], b["offsets"][i + 1]
        fa, fe = full["offsets"][h], full["offsets"][h + 1]
        assert np.array_equal(b["x"][a:e], full["x"][fa:fe])
        assert b["grids"][i] == full["grids"][h]


def _large_catalogue(d, seed=11):
    """60 halos (20 below 30 particles, between large ones), 240k particles:
    DM of two masses in runs, gas of random masses, stars of one mass;
    three halos span the periodic box."""
    rng = np.random.default_rng(seed)
    halos = []
    for h in range(60):
        n = int(rng.integers(5, 29)) if h % 3 == 1 else int(rng.integers(2000, 10000))
        c = rng.uniform(1.0, 99.0, 3)
        if h % 17 == 0:
            c[h % 3] = 99.95
        nd, ng = n * 3 // 5, n // 5
        ns = n - nd - ng

        def put(r):
            p = c + rng.normal(0.0, 0.05, (len(r), 3))
            p %= 100.0
            r["x"], r["y"], r["z"] = p[:, 0], p[:, 1], p[:, 2]
            r["vx"] = rng.normal(0.0, 100.0, len(r))
            return r

        dm = put(np.zeros(nd, dtype=CAT.PM_DTYPE))
        dm["mass"] = np.where((np.arange(nd) // 50) % 2 == 0, 1.0e8, 2.5e8)
        gas = put(np.zeros(ng, dtype=C.GAS))
        gas["mass"] = rng.uniform(1.0e6, 1.0e7, ng).astype(np.float32)
        star = put(np.zeros(ns, dtype=CAT.PM_DTYPE))
        star["mass"] = 1.0e6
        halos.append((dm, gas, np.zeros(0, dtype=np.uint8), star))
    hp, mp = os.path.join(d, "FoF_halo_cat.00002"), os.path.join(d, "FoF_member_particle.00002")
    CAT.write_fof_catalog(hp, mp, [100.0, 0.7, 0.3, 0.048, 0.7, 1.0, 1.0], halos)
    return hp, mp


def test_threaded_batch_equals_small_batches():
    """A batch above the loader's threading size (host threads read their
    halos' records with positional reads) equals the same halos loaded in
    batches of at most 60000 particles (read on the calling thread), field
    by field and record by record, and the reference build's reader where
    it is present; a member file cut short is an error, not a batch."""
    d = tempfile.mkdtemp()
    files = _large_catalogue(d)
    cat = CAT.Catalog(*files)
    big = cat.next_batch()
    assert cat.next_batch() is None
    cat.close()
    n = int(big["offsets"][-1])
    assert n > 200000 and len(big["offsets"]) - 1 == 40
    cat = CAT.Catalog(*files)
    parts = []
    while True:
        b = cat.next_batch(max_particles=60000)
        if b is None:
            break
        assert b["offsets"][-1] <= 60000 or len(b["offsets"]) == 2
        parts.append(b)
    cat.close()
    assert len(parts) > 4
    for k in ("x", "y", "z", "mass", "vx", "vy", "vz", "type", "link02", "records"):
        assert np.array_equal(big[k], np.concatenate([p[k] for p in parts])), k
    for k in ("origin", "grids", "haloq"):
        assert np.array_equal(big[k], np.concatenate([p[k] for p in parts])), k
    sizes = np.concatenate([np.diff(p["offsets"]) for p in parts])
    assert np.array_equal(np.diff(big["offsets"]), sizes)
    from oracle import oracle as O
    import ctypes
    R = O.ref_lib()
    if R is not None:
        R.ref_init(4 << 30)
        cat = CAT.Catalog(*files)
        allh = cat.next_batch(min_np=0)
        cat.close()
        na = int(allh["offsets"][-1])
        x, y, z, m = (np.zeros(na) for _ in range(4))
        ty = np.zeros(na, dtype=np.int32)
        l2 = np.zeros(na, dtype=np.float32)
        off = np.zeros(62, dtype=np.int64)
        org = np.zeros(186)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        R.ref_read_catalog.restype = ctypes.c_longlong
        assert R.ref_read_catalog(files[0].encode(), files[1].encode(), 61,
                                  ctypes.c_longlong(na), p(x), p(y), p(z), p(m), p(ty), p(l2),
                                  p(off), p(org)) == 60
        for k, v in (("x", x), ("y", y), ("z", z), ("mass", m), ("type", ty), ("link02", l2)):
            assert np.array_equal(allh[k], v), k
        assert np.array_equal(allh["origin"].ravel(), org[:180])
    with open(files[1], "r+b") as f:
        f.truncate(os.path.getsize(files[1]) - 1000)
    cat = CAT.Catalog(*files)
    from galaxyfinder_amd._lib import PgalfError
    with pytest.raises(PgalfError):
        cat.next_batch()
    cat.close()
