"""Seeded synthetic FoF catalogue in NewGalFinder's input format, for the
catalogue loader and writer tests.

The files are what opFoF writes and gfind.c reads (gfind.c:25-91,221-241,
363-371): ``FoF_halo_cat.NNNNN`` holds seven floats (size, hubble, omep,
omepb, omeplam, amax, anow) and then one 96-byte HaloQ record per halo
(tree.h:55-60); ``FoF_member_particle.NNNNN`` holds, per halo, its DM, gas,
sink and star particle records in that order (DmType = StarType = PmType of
136 bytes, GasType of 112, SinkType of 168: ramses.h:66-159 under the
reference Makefile's flags -DNCHEM=3, no NBODY).  The record layouts below
are pinned against the reference build (tests/test_catalog.py).

Halos: clumps of a few hundred particles of all four types in a box of
side 100 (cMpc/h), two of them straddling the periodic boundary (one in x,
one in y and z), one smaller than the reference's 30-particle minimum.
"""
import os

import numpy as np

PM = np.dtype({"names": ["x", "y", "z", "vx", "vy", "vz", "mass", "id", "levelp", "family",
                         "tag", "tp", "zp", "chem", "mass0", "birth_d", "partp"],
               "formats": ["<f8"] * 7 + ["<i4", "<i4", "i1", "i1", "<f8", "<f8", ("<f8", 3),
                                         "<f8", "<f8", "<i4"],
               "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 60, 64, 65, 72, 80, 88, 112, 120, 128],
               "itemsize": 136})
GAS = np.dtype({"names": ["x", "y", "z", "cellsize", "vx", "vy", "vz", "den", "temp",
                          "metallicity", "chem", "mass", "potent", "fx", "fy", "fz"],
                "formats": ["<f8"] * 4 + ["<f4"] * 3 + ["<f8", "<f4", "<f4", ("<f4", 3), "<f4",
                                                        "<f8", "<f8", "<f8", "<f8"],
                "offsets": [0, 8, 16, 24, 32, 36, 40, 48, 56, 60, 64, 76, 80, 88, 96, 104],
                "itemsize": 112})
SINK = np.dtype({"names": ["x", "y", "z", "vx", "vy", "vz", "mass", "tbirth", "J", "S", "dM",
                           "E", "id"],
                 "formats": ["<f8"] * 8 + [("<f8", 3)] * 4 + ["<i4"],
                 "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 64, 88, 112, 136, 160],
                 "itemsize": 168})
HALOQ = np.dtype({"names": ["np", "npstar", "npgas", "npdm", "npsink", "x", "y", "z", "mass",
                            "mstar", "mgas", "mdm", "msink", "vx", "vy", "vz"],
                  "formats": ["<u8"] * 5 + ["<f8"] * 3 + ["<f4"] * 8,
                  "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 64, 68, 72, 76, 80, 84, 88, 92],
                  "itemsize": 96})
HALOINFO = np.dtype({"names": ["nsub", "ndm", "nstar", "nsink", "ngas", "npall", "totm", "mdm",
                               "mgas", "msink", "mstar", "x", "y", "z", "vx", "vy", "vz"],
                     "formats": ["<i4"] * 6 + ["<f8"] * 11,
                     "offsets": [0, 4, 8, 12, 16, 20] + [24 + 8 * i for i in range(11)],
                     "itemsize": 112})
SUBINFO = np.dtype({"names": ["npdm", "npgas", "npsink", "npstar", "npall", "totm", "mdm",
                              "mgas", "msink", "mstar", "x", "y", "z", "vx", "vy", "vz"],
                    "formats": ["<i4"] * 5 + ["<f8"] * 11,
                    "offsets": [0, 4, 8, 12, 16] + [24 + 8 * i for i in range(11)],
                    "itemsize": 112})
HEADER = np.array([100.0, 0.7, 0.3, 0.048, 0.7, 1.0, 1.0], dtype="<f4")
TYPE_STAR, TYPE_DM, TYPE_GAS, TYPE_SINK = 1, 2, 3, 4

# (centre, counts dm/gas/sink/star, clump radius)
HALOS = [
    ((30.0, 40.0, 50.0), (300, 120, 2, 80), 0.4),
    ((99.9, 20.0, 60.0), (250, 90, 1, 60), 0.3),     # straddles x = 0 / 100
    ((10.0, 99.95, 0.05), (200, 60, 0, 40), 0.25),   # straddles y and z
    ((70.0, 70.0, 70.0), (15, 5, 0, 4), 0.2),        # fewer than 30 particles
    ((55.0, 5.0, 45.0), (400, 0, 3, 150), 0.5),
]


def _pos(rng, c, n, r):
    p = np.asarray(c) + rng.normal(0.0, r / 2.0, (n, 3))
    return np.mod(p, HEADER[0])


def write(dirname, step=1):
    """Write the two files under dirname; returns their paths."""
    rng = np.random.default_rng(20261204)
    hq = np.zeros(len(HALOS), dtype=HALOQ)
    parts = []
    for h, (c, (ndm, ngas, nsink, nstar), r) in enumerate(HALOS):
        dm = np.zeros(ndm, dtype=PM)
        gas = np.zeros(ngas, dtype=GAS)
        sink = np.zeros(nsink, dtype=SINK)
        star = np.zeros(nstar, dtype=PM)
        for rec, n in ((dm, ndm), (gas, ngas), (sink, nsink), (star, nstar)):
            p = _pos(rng, c, n, r)
            rec["x"], rec["y"], rec["z"] = p[:, 0], p[:, 1], p[:, 2]
            for v in ("vx", "vy", "vz"):
                rec[v] = rng.normal(0.0, 200.0, n)
        dm["mass"] = 1.0e8
        star["mass"] = rng.uniform(1e5, 1e7, nstar)
        gas["mass"] = rng.uniform(1e6, 1e7, ngas).astype(np.float32)
        sink["mass"] = rng.uniform(1e6, 1e8, nsink)
        dm["id"] = np.arange(ndm) + 1000 * h
        star["id"] = np.arange(nstar) + 1000 * h + 500
        star["tp"] = rng.uniform(0, 1, nstar)
        star["chem"] = rng.uniform(0, 0.02, (nstar, 3))
        gas["den"] = rng.uniform(0, 10, ngas)
        gas["temp"] = rng.uniform(1e4, 1e6, ngas).astype(np.float32)
        sink["id"] = np.arange(nsink) + 1000 * h + 900
        n = ndm + ngas + nsink + nstar
        hq[h] = (n, nstar, ngas, ndm, nsink, c[0], c[1], c[2], 0, 0, 0, 0, 0, 0, 0, 0)
        parts.append((dm, gas, sink, star))
    hpath = os.path.join(dirname, "FoF_halo_cat.%05d" % step)
    mpath = os.path.join(dirname, "FoF_member_particle.%05d" % step)
    with open(hpath, "wb") as f:
        f.write(HEADER.tobytes())
        f.write(hq.tobytes())
    with open(mpath, "wb") as f:
        for recs in parts:
            for r in recs:
                f.write(r.tobytes())
    return hpath, mpath


def membership(np_):
    """A seeded subhalo assignment for the writer tests: mpeak and p2halo
    (-9999 for background particles) of a halo of np_ particles."""
    rng = np.random.default_rng(np_)
    mpeak = int(np_ % 4)
    if mpeak == 0:
        return 0, np.full(np_, -9999, dtype=np.int32)
    p2 = rng.integers(-1, mpeak, np_).astype(np.int32)
    p2[p2 < 0] = -9999
    return mpeak, p2
