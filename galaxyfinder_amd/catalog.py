"""FoF catalogue loader and GALFIND writer (pgalf_catalog_* and pgalf_writer_*
of include/pgalf_catalog.h): NewGalFinder's input catalogue read in batches
of halos, localised as subhalo_den does (Dump2BP, mklocalize) with the TSC
grid of lagFindStellarCore per halo, ready for ``density.Batch``; and
write_data's output files (gfind.c:700-732).  Host code; no GPU needed
except to run the batch."""
import ctypes

import numpy as np

from . import _lib
from ._lib import pgalf_grid_t

UNWRAP_REFERENCE, UNWRAP_SORTED = 0, 1
MIN_HALO_NP = 30              # gfind.c:388
GAUSSIAN_SMOOTHING_LENGTH = 0.008  # params.h:104 (the RG of every halo)
NOT_HALO_MEMBER = -9999       # defs.h:29

HALOQ_DTYPE = np.dtype({"names": ["np", "npstar", "npgas", "npdm", "npsink", "x", "y", "z",
                                  "mass", "mstar", "mgas", "mdm", "msink", "vx", "vy", "vz"],
                        "formats": ["<u8"] * 5 + ["<f8"] * 3 + ["<f4"] * 8,
                        "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 64, 68, 72, 76, 80, 84, 88,
                                    92],
                        "itemsize": 96})
GRID_DTYPE = np.dtype({"names": ["nx", "ny", "nz", "xmin", "ymin", "zmin", "cellsize"],
                       "formats": ["<i4"] * 3 + ["<f8"] * 4,
                       "offsets": [0, 4, 8, 16, 24, 32, 40], "itemsize": 48})


PM_DTYPE = np.dtype({"names": ["x", "y", "z", "vx", "vy", "vz", "mass", "id", "levelp",
                               "family", "tag", "tp", "zp", "chem", "mass0", "birth_d",
                               "partp"],
                     "formats": ["<f8"] * 7 + ["<i4", "<i4", "i1", "i1", "<f8", "<f8",
                                               ("<f8", 3), "<f8", "<f8", "<i4"],
                     "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 60, 64, 65, 72, 80, 88, 112,
                                 120, 128],
                     "itemsize": 136})  # DmType = StarType (ramses.h:66-92)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def write_fof_catalog(halo_cat, members, header, halos):
    """Writes a catalogue in the FoF format the loader reads: header (7
    floats: size, hubble, omep, omepb, omeplam, amax, anow) and, per halo,
    a HaloQ record and its records.  halos: list of (dm, gas, sink, star)
    structured arrays (PM_DTYPE for dm and star; gas and sink of the
    reference's GasType / SinkType layouts, or empty)."""
    hq = np.zeros(len(halos), dtype=HALOQ_DTYPE)
    for h, (dm, gas, sink, star) in enumerate(halos):
        n = len(dm) + len(gas) + len(sink) + len(star)
        hq[h]["np"], hq[h]["npstar"], hq[h]["npgas"] = n, len(star), len(gas)
        hq[h]["npdm"], hq[h]["npsink"] = len(dm), len(sink)
        if len(dm):
            hq[h]["x"], hq[h]["y"], hq[h]["z"] = dm["x"][0], dm["y"][0], dm["z"][0]
    with open(halo_cat, "wb") as f:
        f.write(np.asarray(header, dtype="<f4").tobytes())
        f.write(hq.tobytes())
    with open(members, "wb") as f:
        for recs in halos:
            for r in recs:
                if len(r):
                    f.write(r.tobytes())


class Catalog:
    """Reader of FoF_halo_cat.* / FoF_member_particle.* (pgalf_catalog_*)."""

    def __init__(self, halo_cat, members, unwrap=UNWRAP_REFERENCE):
        self._lib = _lib.load()
        self._c = self._lib.pgalf_catalog_open(halo_cat.encode(), members.encode(), int(unwrap))
        if not self._c:
            _lib.check()
        h = (ctypes.c_float * 7)()
        _lib.check(self._lib.pgalf_catalog_header(self._c, h))
        self.header = dict(zip(("size", "hubble", "omep", "omepb", "omeplam", "amax", "anow"),
                               list(h)))

    def next_batch(self, max_halos=1 << 30, max_particles=1 << 40, min_np=MIN_HALO_NP,
                   copy_records=True):
        """The next batch of halos as a dict of numpy arrays (None at the
        end of the catalogue): x, y, z (localised, FP64), mass, type, link02,
        vx, vy, vz per particle; offsets, record_offsets (bytes), haloq,
        origin (nhalo x 3), grids (pgalf_grid_t fields) per halo; records
        (the member-file bytes of the batch; with copy_records=False a view
        of the library's buffer, valid until the next call or close)."""
        npart = ctypes.c_longlong()
        nh = self._lib.pgalf_catalog_next(self._c, int(max_halos), int(max_particles),
                                          int(min_np), ctypes.byref(npart))
        if nh < 0:
            _lib.check()
        if nh == 0:
            return None
        n = npart.value
        out = {k: np.empty(n) for k in ("x", "y", "z", "mass", "vx", "vy", "vz")}
        out["type"] = np.empty(n, dtype=np.int32)
        out["link02"] = np.empty(n, dtype=np.float32)
        out["offsets"] = np.empty(nh + 1, dtype=np.int64)
        out["record_offsets"] = np.empty(nh + 1, dtype=np.int64)
        out["haloq"] = np.empty(nh, dtype=HALOQ_DTYPE)
        out["origin"] = np.empty((nh, 3))
        out["grids"] = np.empty(nh, dtype=GRID_DTYPE)
        _lib.check(self._lib.pgalf_catalog_fetch(
            self._c, _p(out["x"]), _p(out["y"]), _p(out["z"]), _p(out["mass"]), _p(out["type"]),
            _p(out["link02"]), _p(out["vx"]), _p(out["vy"]), _p(out["vz"]), _p(out["offsets"]),
            _p(out["record_offsets"]), _p(out["haloq"]), _p(out["origin"]), _p(out["grids"])))
        nb = ctypes.c_longlong()
        ptr = self._lib.pgalf_catalog_records(self._c, ctypes.byref(nb))
        if nb.value:
            view = np.ctypeslib.as_array((ctypes.c_ubyte * nb.value).from_address(ptr))
            out["records"] = view.copy() if copy_records else view
        else:
            out["records"] = np.zeros(0, dtype=np.uint8)
        return out

    @staticmethod
    def grid_tuples(batch):
        g = batch["grids"]
        return [(int(r["nx"]), int(r["ny"]), int(r["nz"]), float(r["xmin"]), float(r["ymin"]),
                 float(r["zmin"]), float(r["cellsize"])) for r in g]

    def close(self):
        if getattr(self, "_c", None):
            self._lib.pgalf_catalog_close(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Writer:
    """GALFIND.DATA / GALCATALOG.LIST / background writer (pgalf_writer_*)."""

    def __init__(self, galfind_data, catalog_list, background=None):
        self._lib = _lib.load()
        self._w = self._lib.pgalf_writer_open(galfind_data.encode(), catalog_list.encode(),
                                              background.encode() if background else None)
        if not self._w:
            _lib.check()

    def halo(self, types, records, p2halo, mpeak):
        t = np.ascontiguousarray(types, dtype=np.int32)
        r = np.ascontiguousarray(records, dtype=np.uint8)
        p = np.ascontiguousarray(p2halo, dtype=np.int32)
        if len(p) != len(t):
            raise ValueError("p2halo and types differ in length")
        _lib.check(self._lib.pgalf_writer_halo(self._w, len(t), _p(t), _p(r), _p(p), int(mpeak)))

    def close(self):
        if getattr(self, "_w", None):
            _lib.check(self._lib.pgalf_writer_close(self._w))
            self._w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
