"""Python mirror of NewGalFinder's density-stage interface, backed by the HIP
library (libpgalf_density.so).

Host functions keep the reference's names and argument meaning
(tsc_omp2.c:40,166; gsmooth.c:14; nnost.c:560,632) with numpy arrays in
place of C pointers; outputs are written into the caller's arrays as the C
functions do.  Failures raise ``PgalfError`` (the C library reports a
status instead of calling exit()).

Device functions (``dev_*``) take float32 torch tensors resident on the GPU
and run on torch's current stream.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import PgalfError, pgalf_grid_t, pgalf_shells_t

TYPE_STAR = 1
TYPE_DM = 2
TYPE_GAS = 3
TYPE_SINK = 4
TYPE_ALL = 5

# SimpleBasicParticleType (tree.h:155-162), 88 bytes.
PARTICLE_DTYPE = np.dtype({
    "names": ["type", "mass", "x", "y", "z", "vx", "vy", "vz", "link02", "indx", "bp"],
    "formats": ["<u4", "<f8", "<f8", "<f8", "<f8", "<f8", "<f8", "<f8", "<f8", "<i4", "<u8"],
    "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 64, 72, 80],
    "itemsize": 88,
})


# Coretype (tree.h:204-213), 56 bytes.
CORE_DTYPE = np.dtype({
    "names": ["peak", "nummem", "numstar", "starmass", "coredensity", "cx", "cy", "cz",
              "cvx", "cvy", "cvz", "Rtidal", "density", "flag"],
    "formats": ["<i4", "<i4", "<i4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4", "<f4",
                "<f4", "<f4", "u1"],
    "offsets": [0, 4, 8, 12, 16, 20, 24, 28, 32, 36, 40, 44, 48, 52],
    "itemsize": 56,
})
PEAKTHRESHOLD = 2000.0  # params.h:19
NCELLBUFF = 10          # params.h:115


def make_particles(pos, mass=None, types=None):
    """Pack positions (n,3), masses and types into reference particle records."""
    pos = np.asarray(pos, dtype=np.float64)
    n = len(pos)
    bp = np.zeros(n, dtype=PARTICLE_DTYPE)
    bp["x"], bp["y"], bp["z"] = pos[:, 0], pos[:, 1], pos[:, 2]
    bp["mass"] = 1.0 if mass is None else np.asarray(mass, dtype=np.float64)
    bp["type"] = TYPE_STAR if types is None else np.asarray(types, dtype=np.uint32)
    bp["indx"] = np.arange(n, dtype=np.int32)
    return bp


def padded_shape(nx, ny, nz):
    """(nz, ny, mx) of the in-place r2c grid, mx = 2(nx/2+1) (tsc_omp2.c:198)."""
    return (nz, ny, 2 * (nx // 2 + 1))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _need(a, dtype, n, what):
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous:
        raise TypeError("%s must be a C-contiguous %s array" % (what, np.dtype(dtype).name))
    if a.size < n:
        raise ValueError("%s holds %d elements, needs %d" % (what, a.size, n))


def assign_density_TSC(bp, np_, den, nx, ny, nz, Xmin, Ymin, Zmin, cellsize, ptype):
    """TSC deposit into den (float32, mx*ny*nz) -- tsc_omp2.c:166."""
    lib = _lib.load()
    _need(bp, PARTICLE_DTYPE, np_, "bp")
    _need(den, np.float32, 2 * (nx // 2 + 1) * ny * nz, "den")
    lib.assign_density_TSC(_ptr(bp), int(np_), _ptr(den), nx, ny, nz, Xmin, Ymin, Zmin,
                           cellsize, ptype)
    _lib.check()


def gaussian_Smoothing(denGrid, nx, ny, nz, cellsize, RG):
    """In-place FFT Gaussian smoothing of the padded grid -- gsmooth.c:14."""
    lib = _lib.load()
    _need(denGrid, np.float32, 2 * (nx // 2 + 1) * ny * nz, "denGrid")
    lib.gaussian_Smoothing(_ptr(denGrid), nx, ny, nz, cellsize, RG)
    _lib.check()


def findDen(bp, np_, densph, xmin, ymin, zmin, cellsize, den, nx, ny, nz):
    """TSC read-back into densph (float32, np) -- tsc_omp2.c:40."""
    lib = _lib.load()
    _need(bp, PARTICLE_DTYPE, np_, "bp")
    _need(densph, np.float32, np_, "densph")
    _need(den, np.float32, 2 * (nx // 2 + 1) * ny * nz, "den")
    lib.findDen(_ptr(bp), int(np_), _ptr(densph), xmin, ymin, zmin, cellsize, _ptr(den),
                nx, ny, nz)
    _lib.check()


def findsphdensity(bp, np_, nearindex, Numnear, densph):
    """Exact Numnear-NN lists + unweighted W4 density -- nnost.c:560."""
    lib = _lib.load()
    _need(bp, PARTICLE_DTYPE, np_, "bp")
    _need(nearindex, np.int64, np_ * Numnear, "nearindex")
    _need(densph, np.float32, np_, "densph")
    lib.findsphdensity(_ptr(bp), int(np_), _ptr(nearindex), int(Numnear), _ptr(densph))
    _lib.check()


def starfindsphdensity(bp, np_, nearindex, Numnear, densph):
    """Star-weighted W4 density (10 nearest stars) + all-particle lists --
    nnost.c:632."""
    lib = _lib.load()
    _need(bp, PARTICLE_DTYPE, np_, "bp")
    _need(nearindex, np.int64, np_ * Numnear, "nearindex")
    _need(densph, np.float32, np_, "densph")
    lib.starfindsphdensity(_ptr(bp), int(np_), _ptr(nearindex), int(Numnear), _ptr(densph))
    _lib.check()


def near_lists(bp, np_, Numnear, nearindex):
    """All-particle Numnear-NN lists of lagFindStellarCore (nnost.c:1298-1536)."""
    lib = _lib.load()
    _need(bp, PARTICLE_DTYPE, np_, "bp")
    _need(nearindex, np.int64, np_ * Numnear, "nearindex")
    _lib.check(lib.pgalf_near_lists(_ptr(bp), int(np_), int(Numnear), _ptr(nearindex)))


def lag_density_stage(bp, Numnear, ptype, want_grid=False):
    """The density stage of lagFindStellarCore (nnost.c:1030-1096,
    1298-1536): returns (densph, nearindex[np,Numnear], grid or None,
    (nx,ny,nz), (xmin,ymin,zmin))."""
    lib = _lib.load()
    n = len(bp)
    _need(bp, PARTICLE_DTYPE, n, "bp")
    densph = np.zeros(n, dtype=np.float32)
    nbr = np.zeros((n, Numnear), dtype=np.int64)
    gptr = ctypes.POINTER(ctypes.c_float)()
    nxyz = (ctypes.c_int * 3)()
    x0 = (ctypes.c_double * 3)()
    rc = lib.pgalf_lag_density_stage(_ptr(bp), n, int(Numnear), int(ptype), _ptr(densph),
                                     _ptr(nbr), ctypes.byref(gptr) if want_grid else None,
                                     nxyz, x0)
    grid = None
    if want_grid and gptr:
        nx, ny, nz = nxyz[0], nxyz[1], nxyz[2]
        count = 2 * (nx // 2 + 1) * ny * nz
        grid = np.ctypeslib.as_array(gptr, shape=(count,)).copy().reshape(padded_shape(nx, ny, nz))
        libc = ctypes.CDLL(None)
        libc.free(gptr)
    _lib.check(rc)
    return densph, nbr, grid, tuple(nxyz), tuple(x0)


def _cores(fn, maxcore, *args):
    n = ctypes.c_int(0)
    core = np.zeros(max(int(maxcore), 1), dtype=CORE_DTYPE)
    rc = fn(*args, _ptr(core), int(maxcore), ctypes.byref(n))
    _lib.check(rc)
    return core[:n.value]


def grid_cores(bp, densph, denGrid, nx, ny, nz, xmin, ymin, zmin, cellsize, ptype,
               threshold=PEAKTHRESHOLD, nbuff=NCELLBUFF, maxcore=None):
    """Grid cores of lagFindStellarCore (nnost.c:1115-1271) as an array of
    Coretype records in (k, j, i) node order."""
    lib = _lib.load()
    n = len(bp)
    _need(bp, PARTICLE_DTYPE, n, "bp")
    _need(densph, np.float32, n, "densph")
    _need(denGrid, np.float32, 2 * (nx // 2 + 1) * ny * nz, "denGrid")
    maxcore = max(1, nx * ny * nz // 8) if maxcore is None else maxcore
    return _cores(lib.pgalf_grid_cores, maxcore, _ptr(bp), n, _ptr(densph), _ptr(denGrid),
                  nx, ny, nz, xmin, ymin, zmin, cellsize, int(ptype), float(threshold),
                  int(nbuff))


def finddenpeak(den, numneigh, neighbor, jflag, bp, maxcore=None):
    """finddenpeak (subhaloden.mod6.c:416-492) as an array of Coretype
    records in index order."""
    lib = _lib.load()
    n = len(bp)
    _need(bp, PARTICLE_DTYPE, n, "bp")
    _need(den, np.float32, n, "den")
    _need(neighbor, np.int64, n * numneigh, "neighbor")
    maxcore = max(1, n) if maxcore is None else maxcore
    return _cores(lib.pgalf_finddenpeak, maxcore, _ptr(den), int(numneigh), _ptr(neighbor), n,
                  int(jflag), _ptr(bp))


MERGINGPEAKLENGTH = float(np.float32(4.0e-3))  # params.h:49 (PGALF_MERGINGPEAKLENGTH)


def MergingPeak(bp, np_, core, numcore, iflag=0):
    """MergingPeak (subhaloden.mod6.c:341-415) on a Coretype record array, in
    place, as the reference: returns the new core count; core[:count] are the
    kept cores, largest key first (density for iflag 0, numstar for 1)."""
    lib = _lib.load()
    _need(bp, PARTICLE_DTYPE, np_, "bp")
    _need(core, CORE_DTYPE, numcore, "core")
    n = lib.pgalf_merging_peak(_ptr(bp), int(np_), _ptr(core), int(numcore), int(iflag))
    if n < 0:
        _lib.check()
    return n


# Watershed of subhalo_den (subhaloden.mod6.c:1167-1475, 572-678, 2995-3024).
# The reference's working particles `wp` are three arrays here: den
# (wp[].den), haloid (wp[].haloid) and flag (wp[].flag).
WP_PEAK, WP_VISITED, WP_CORE, WP_SHELL, WP_BOUND, WP_REMAINING = 1, 2, 4, 8, 16, 32
NOT_HALO_MEMBER = -9999  # defs.h:29
MAX_SHELL_DIVIDE = 15


def FindCoreDensity(bp, np_, neighbor, NumNeighbor, core, numcore, den, haloid, flag):
    """FindCoreDensity (subhaloden.mod6.c:1167-1475) with wp as den, haloid
    (int32, out) and flag (uint8, out): fills core[:numcore]'s coredensity,
    nummem, numstar and starmass in place; returns numcore."""
    lib = _lib.load()
    _need(bp, PARTICLE_DTYPE, np_, "bp")
    _need(neighbor, np.int64, np_ * NumNeighbor, "neighbor")
    _need(core, CORE_DTYPE, numcore, "core")
    _need(den, np.float32, np_, "den")
    _need(haloid, np.int32, np_, "haloid")
    _need(flag, np.uint8, np_, "flag")
    n = lib.pgalf_find_core_density(_ptr(bp), int(np_), _ptr(neighbor), int(NumNeighbor),
                                    _ptr(core), int(numcore), _ptr(den), _ptr(haloid),
                                    _ptr(flag))
    if n < 0:
        _lib.check()
    return n


def shell_thresholds(den, np_, core, numcore):
    """The shell thresholds of subhalo_den (subhaloden.mod6.c:2995-3024),
    highest first (float32 array)."""
    lib = _lib.load()
    _need(den, np.float32, np_, "den")
    _need(core, CORE_DTYPE, numcore, "core")
    thr = np.zeros(MAX_SHELL_DIVIDE, dtype=np.float32)
    n = lib.pgalf_shell_thresholds(_ptr(den), int(np_), _ptr(core), int(numcore), _ptr(thr))
    if n < 0:
        _lib.check()
    return thr[:n].copy()


def _shell_lists(n, part_ptr, part, core_ptr, cores):
    return [(part[part_ptr[i]:part_ptr[i + 1]].copy(), cores[core_ptr[i]:core_ptr[i + 1]].copy())
            for i in range(n)]


def GetShellParticlesPeaks(np_, neighbor, NumNeighbor, core, numcore, den, haloid, flag, thr,
                           maxcores=None):
    """GetShellParticlesPeaks (subhaloden.mod6.c:612-678) for every threshold
    of thr in order, then SaveRemainingParticles2LastShell (:572-611); flag
    and haloid are updated in place.  Returns the shells as a list of
    (particles, cores) arrays (SHELL2P / SHELL2C)."""
    lib = _lib.load()
    _need(neighbor, np.int64, np_ * NumNeighbor, "neighbor")
    _need(core, CORE_DTYPE, numcore, "core")
    _need(den, np.float32, np_, "den")
    _need(haloid, np.int32, np_, "haloid")
    _need(flag, np.uint8, np_, "flag")
    thr = np.ascontiguousarray(thr, dtype=np.float32)
    maxshell = numcore * max(len(thr), 1) + 1
    maxcores = 16 * numcore * (len(thr) + 1) + 16 if maxcores is None else maxcores
    while True:
        pp = np.zeros(maxshell + 1, np.int64)
        cp = np.zeros(maxshell + 1, np.int64)
        part = np.zeros(max(np_, 1), np.int32)
        cores = np.zeros(maxcores, np.int32)
        sh = pgalf_shells_t(_ptr(pp), _ptr(part), _ptr(cp), _ptr(cores), maxshell, maxcores, 0)
        fl, hid = flag.copy(), haloid.copy()
        rc = lib.pgalf_shell_particles_peaks(int(np_), _ptr(neighbor), int(NumNeighbor),
                                             _ptr(core), int(numcore), _ptr(den), _ptr(hid),
                                             _ptr(fl), _ptr(thr), len(thr), ctypes.byref(sh))
        if rc == 6 and maxcores < 1024 * (numcore + 1) * (len(thr) + 1):  # core entries overflow
            maxcores *= 4
            continue
        _lib.check(rc)
        flag[:] = fl
        haloid[:] = hid
        return _shell_lists(sh.nshell, pp, part, cp, cores)


# ---------------------------------------------------------------------------
# device-resident API (torch tensors on the GPU)
# ---------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_dev(t, dtype, n, what):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != dtype or \
            not t.is_contiguous() or t.numel() < n:
        raise TypeError("%s must be a contiguous %s CUDA tensor with >= %d elements"
                        % (what, dtype, n))


def grid_struct(nx, ny, nz, xmin, ymin, zmin, cellsize):
    return pgalf_grid_t(int(nx), int(ny), int(nz), float(xmin), float(ymin), float(zmin),
                        float(cellsize))


def dev_tsc(x, y, z, m, grid, den, ptype=TYPE_ALL, types=None):
    torch = _torch()
    n = x.numel()
    for t, w in ((x, "x"), (y, "y"), (z, "z"), (m, "m")):
        _check_dev(t, torch.float32, n, w)
    if types is not None:
        _check_dev(types, torch.int32, n, "types")
    _check_dev(den, torch.float32, 2 * (grid.nx // 2 + 1) * grid.ny * grid.nz, "den")
    lib = _lib.load()
    _lib.check(lib.pgalf_dev_tsc(_dptr(x), _dptr(y), _dptr(z), _dptr(m), _dptr(types), n,
                                 int(ptype), ctypes.byref(grid), _dptr(den), _stream()))


def dev_smooth(den, grid, RG):
    torch = _torch()
    _check_dev(den, torch.float32, 2 * (grid.nx // 2 + 1) * grid.ny * grid.nz, "den")
    lib = _lib.load()
    _lib.check(lib.pgalf_dev_smooth(_dptr(den), ctypes.byref(grid), float(RG), _stream()))


def _grid_floats(grid):
    return 2 * (grid.nx // 2 + 1) * grid.ny * grid.nz


def dev_findden(x, y, z, grid, den, out):
    torch = _torch()
    n = x.numel()
    for t, w in ((x, "x"), (y, "y"), (z, "z"), (out, "out")):
        _check_dev(t, torch.float32, n, w)
    _check_dev(den, torch.float32, _grid_floats(grid), "den")
    lib = _lib.load()
    _lib.check(lib.pgalf_dev_findden(_dptr(x), _dptr(y), _dptr(z), n, ctypes.byref(grid),
                                     _dptr(den), _dptr(out), _stream()))


def dev_knn(x, y, z, k, m=None, nbr=None, rho=None, rho_mode=0, types=None,
            cand_ptype=TYPE_ALL):
    torch = _torch()
    n = x.numel()
    for t, w in ((x, "x"), (y, "y"), (z, "z")):
        _check_dev(t, torch.float32, n, w)
    if nbr is not None:
        _check_dev(nbr, torch.int64, n * k, "nbr")
    if rho is not None:
        _check_dev(rho, torch.float32, n, "rho")
    if rho_mode == 2 or m is not None:
        _check_dev(m, torch.float32, n, "m")
    if types is not None:
        _check_dev(types, torch.int32, n, "types")
    lib = _lib.load()
    _lib.check(lib.pgalf_dev_knn(_dptr(x), _dptr(y), _dptr(z), _dptr(m), _dptr(types), n,
                                 int(cand_ptype), int(k), _dptr(nbr), _dptr(rho),
                                 int(rho_mode), _stream()))


def dev_density_stage(x, y, z, m, grid, RG, k, den, rho_tsc=None, nbr=None, rho_knn=None):
    """deposit -> smooth -> read-back -> kNN lists + W4 density, one call."""
    torch = _torch()
    n = x.numel()
    for t, w in ((x, "x"), (y, "y"), (z, "z"), (m, "m")):
        _check_dev(t, torch.float32, n, w)
    _check_dev(den, torch.float32, _grid_floats(grid), "den")
    for t, w in ((rho_tsc, "rho_tsc"), (rho_knn, "rho_knn")):
        if t is not None:
            _check_dev(t, torch.float32, n, w)
    if nbr is not None:
        _check_dev(nbr, torch.int64, n * k, "nbr")
    lib = _lib.load()
    _lib.check(lib.pgalf_dev_density_stage(_dptr(x), _dptr(y), _dptr(z), _dptr(m), n,
                                           ctypes.byref(grid), float(RG), int(k), _dptr(den),
                                           _dptr(rho_tsc), _dptr(nbr), _dptr(rho_knn),
                                           _stream()))


def dev_density_stage_backward(x, y, z, m, grid, RG, k, den, nbr, rho_mode, g_tsc=None,
                               g_knn=None):
    """Adjoint of the device stage: returns (dL/dx, dL/dy, dL/dz, dL/dm)
    float32 tensors from the cotangents of the read-back density (g_tsc)
    and of the W4 kNN density (g_knn, rho_mode 1 unweighted / 2 weighted)."""
    torch = _torch()
    n = x.numel()
    for t, w in ((x, "x"), (y, "y"), (z, "z"), (m, "m")):
        _check_dev(t, torch.float32, n, w)
    if g_tsc is not None:
        _check_dev(g_tsc, torch.float32, n, "g_tsc")
        _check_dev(den, torch.float32, 2 * (grid.nx // 2 + 1) * grid.ny * grid.nz, "den")
    if g_knn is not None:
        _check_dev(g_knn, torch.float32, n, "g_knn")
        _check_dev(nbr, torch.int64, n * k, "nbr")
    out = [torch.empty(n, dtype=torch.float32, device=x.device) for _ in range(4)]
    lib = _lib.load()
    _lib.check(lib.pgalf_dev_density_stage_backward(
        _dptr(x), _dptr(y), _dptr(z), _dptr(m), n, ctypes.byref(grid), float(RG), int(k),
        _dptr(den) if g_tsc is not None else None, _dptr(nbr) if g_knn is not None else None,
        int(rho_mode), _dptr(g_tsc), _dptr(g_knn), _dptr(out[0]), _dptr(out[1]),
        _dptr(out[2]), _dptr(out[3]), _stream()))
    return tuple(out)


class Batch:
    """A catalogue of halos run through the density stage together
    (pgalf_batch_* of the C-ABI) -- the batched counterpart of calling the
    stage once per halo as the reference's workers do (gfind.c:394-549).

    offsets: int64 [nhalo+1], halo h owns particles [offsets[h], offsets[h+1])
    of the concatenated arrays (offsets[0] == 0).  grids: nhalo grid specs
    (pgalf_grid_t or (nx, ny, nz, xmin, ymin, zmin, cellsize) tuples).  RG:
    per-halo smoothing lengths in box units.  k: neighbours per particle.
    The grid buffer holds every halo's padded grid; halo h's starts at float
    ``grid_offsets[h]`` and the buffer has ``grid_offsets[-1]`` floats.
    Neighbour indices are local to each halo.
    """

    def __init__(self, offsets, grids, RG, k):
        lib = _lib.load()
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        nhalo = len(off) - 1
        garr = (pgalf_grid_t * nhalo)()
        for h, g in enumerate(grids):
            if isinstance(g, pgalf_grid_t):
                garr[h] = g
            else:
                garr[h] = grid_struct(*g)
        rg = np.ascontiguousarray(np.broadcast_to(np.asarray(RG, dtype=np.float32), (nhalo,)))
        self.nhalo, self.k = nhalo, int(k)
        self.offsets = off
        self._b = lib.pgalf_batch_create(nhalo, _ptr(off), garr, _ptr(rg), int(k))
        if not self._b:
            _lib.check()
            raise PgalfError(-1, "pgalf_batch_create failed")
        self.grid_offsets = np.array([lib.pgalf_batch_grid_offset(self._b, h)
                                      for h in range(nhalo + 1)], dtype=np.int64)
        self.grids = [(garr[h].nx, garr[h].ny, garr[h].nz) for h in range(nhalo)]

    @property
    def grid_floats(self):
        return int(self.grid_offsets[-1])

    def grid_view(self, den, h):
        """Halo h's padded (nz, ny, mx) grid inside the batch grid buffer."""
        nx, ny, nz = self.grids[h]
        a, b = self.grid_offsets[h], self.grid_offsets[h + 1]
        return den[a:b].reshape(padded_shape(nx, ny, nz))

    def run(self, x, y, z, m, den, rho_tsc=None, nbr=None, rho_knn=None, graph=False):
        torch = _torch()
        n = int(self.offsets[-1])
        for t, w in ((x, "x"), (y, "y"), (z, "z"), (m, "m")):
            _check_dev(t, torch.float32, n, w)
        _check_dev(den, torch.float32, self.grid_floats, "den")
        if rho_tsc is not None:
            _check_dev(rho_tsc, torch.float32, n, "rho_tsc")
        if nbr is not None:
            _check_dev(nbr, torch.int64, n * self.k, "nbr")
        if rho_knn is not None:
            _check_dev(rho_knn, torch.float32, n, "rho_knn")
        lib = _lib.load()
        _lib.check(lib.pgalf_batch_run(self._b, _dptr(x), _dptr(y), _dptr(z), _dptr(m),
                                       _dptr(den), _dptr(rho_tsc), _dptr(nbr), _dptr(rho_knn),
                                       1 if graph else 0, _stream()))

    def run_grad(self, x, y, z, m, w_tsc, w_knn, rho_mode=2, climb_iters=5, climb_step=0.5,
                 graph=False, out=None):
        """Forward, cotangents of sum w_tsc log rho_tsc + sum w_knn log rho_knn,
        adjoint and peak climbing in one launch sequence (pgalf_batch_run_grad).
        Returns a dict of device tensors: den, rho_tsc, nbr, rho_knn, gx, gy,
        gz, gm, cx, cy, cz, peak (``out`` may supply them to reuse buffers)."""
        torch = _torch()
        n = int(self.offsets[-1])
        dev = x.device
        for t, w in ((x, "x"), (y, "y"), (z, "z"), (m, "m")):
            _check_dev(t, torch.float32, n, w)
        for t, w in ((w_tsc, "w_tsc"), (w_knn, "w_knn")):
            if t is not None:
                _check_dev(t, torch.float32, n, w)
        if out is None and graph:
            # a graph is captured per set of pointers: reuse one output set
            out = getattr(self, "_grad_out", None)
        if out is None:
            f32 = dict(dtype=torch.float32, device=dev)
            out = dict(den=torch.empty(self.grid_floats, **f32),
                       rho_tsc=torch.empty(n, **f32), rho_knn=torch.empty(n, **f32),
                       nbr=torch.empty((n, self.k), dtype=torch.int64, device=dev),
                       gx=torch.empty(n, **f32), gy=torch.empty(n, **f32),
                       gz=torch.empty(n, **f32), gm=torch.empty(n, **f32),
                       cx=torch.empty(n, **f32), cy=torch.empty(n, **f32),
                       cz=torch.empty(n, **f32),
                       peak=torch.empty(n, dtype=torch.int32, device=dev))
            if graph:
                self._grad_out = out
        shapes = dict(den=(torch.float32, self.grid_floats), rho_tsc=(torch.float32, n),
                      rho_knn=(torch.float32, n), nbr=(torch.int64, n * self.k),
                      gx=(torch.float32, n), gy=(torch.float32, n), gz=(torch.float32, n),
                      gm=(torch.float32, n), cx=(torch.float32, n), cy=(torch.float32, n),
                      cz=(torch.float32, n), peak=(torch.int32, n))
        for key, (dt, cnt) in shapes.items():
            _check_dev(out.get(key), dt, cnt, "out[%r]" % key)
        o = out
        lib = _lib.load()
        _lib.check(lib.pgalf_batch_run_grad(
            self._b, _dptr(x), _dptr(y), _dptr(z), _dptr(m), _dptr(w_tsc), _dptr(w_knn),
            int(rho_mode), _dptr(o["den"]), _dptr(o["rho_tsc"]), _dptr(o["nbr"]),
            _dptr(o["rho_knn"]), _dptr(o["gx"]), _dptr(o["gy"]), _dptr(o["gz"]),
            _dptr(o["gm"]), int(climb_iters), float(climb_step), _dptr(o["cx"]),
            _dptr(o["cy"]), _dptr(o["cz"]), _dptr(o["peak"]), 1 if graph else 0, _stream()))
        return out

    def workspace_bytes(self, reserved=False):
        """Largest device scratch any run of this batch has held at once (or,
        with reserved=True, the bytes its arena holds now)."""
        return int(_lib.load().pgalf_batch_workspace_bytes(self._b, 1 if reserved else 0))

    def status(self):
        """Synchronise torch's current stream; raise on kernel-detected errors."""
        lib = _lib.load()
        _lib.check(lib.pgalf_batch_status(self._b, _stream()))

    def close(self):
        if getattr(self, "_b", None):
            _lib.load().pgalf_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dev_peak_climb(x, y, z, grid, den, niter=5, step=0.5):
    """Peak climbing on the smoothed grid den (pgalf_dev_peak_climb): returns
    (climbed x, y, z, peak flat index int32)."""
    torch = _torch()
    n = x.numel()
    for t, w in ((x, "x"), (y, "y"), (z, "z")):
        _check_dev(t, torch.float32, n, w)
    _check_dev(den, torch.float32, _grid_floats(grid), "den")
    ox, oy, oz = (torch.empty(n, dtype=torch.float32, device=x.device) for _ in range(3))
    peak = torch.empty(n, dtype=torch.int32, device=x.device)
    lib = _lib.load()
    _lib.check(lib.pgalf_dev_peak_climb(_dptr(x), _dptr(y), _dptr(z), n, ctypes.byref(grid),
                                        _dptr(den), int(niter), float(step), _dptr(ox),
                                        _dptr(oy), _dptr(oz), _dptr(peak), _stream()))
    return ox, oy, oz, peak


def _dev_cores(fn, maxcore, *args):
    torch = _torch()
    core = torch.empty(max(int(maxcore), 1), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    _lib.check(fn(*args, _dptr(core), int(maxcore), _dptr(cnt), _stream()))
    return core, cnt


def dev_grid_cores(x, y, z, densph, grid, den, ptype=TYPE_ALL, types=None,
                   threshold=PEAKTHRESHOLD, nbuff=NCELLBUFF, maxcore=None):
    """Grid cores on device (pgalf_dev_grid_cores): returns (core particle
    indices int32 tensor of capacity maxcore, count int32 tensor [1]); the
    first min(count, maxcore) entries are the cores in (k, j, i) order."""
    torch = _torch()
    n = x.numel()
    for t, w in ((x, "x"), (y, "y"), (z, "z"), (densph, "densph")):
        _check_dev(t, torch.float32, n, w)
    if types is not None:
        _check_dev(types, torch.int32, n, "types")
    _check_dev(den, torch.float32, _grid_floats(grid), "den")
    maxcore = max(1, grid.nx * grid.ny * grid.nz // 8) if maxcore is None else maxcore
    lib = _lib.load()
    return _dev_cores(lib.pgalf_dev_grid_cores, maxcore, _dptr(x), _dptr(y), _dptr(z),
                      _dptr(types), n, int(ptype), _dptr(densph), ctypes.byref(grid), _dptr(den),
                      float(threshold), int(nbuff))


def dev_finddenpeak(den, nbr, types=None, jflag=0, threshold=PEAKTHRESHOLD, maxcore=None):
    """finddenpeak on device (pgalf_dev_finddenpeak): (core indices, count)
    as dev_grid_cores, cores in index order."""
    torch = _torch()
    n = den.numel()
    k = nbr.shape[1]
    _check_dev(den, torch.float32, n, "den")
    _check_dev(nbr, torch.int64, n * k, "nbr")
    if types is not None:
        _check_dev(types, torch.int32, n, "types")
    maxcore = max(1, n) if maxcore is None else maxcore
    lib = _lib.load()
    return _dev_cores(lib.pgalf_dev_finddenpeak, maxcore, _dptr(den), int(k), _dptr(nbr), n,
                      _dptr(types), int(jflag), float(threshold))


def dev_merging_peak(x, y, z, peak, key, link=MERGINGPEAKLENGTH, order=False):
    """MergingPeak on device (pgalf_dev_merging_peak): cores at particles
    peak (int32), ranked by key (float32), largest first.  Returns (keep,
    count): the input indices of the kept cores in rank order in keep[:count]
    (int32 tensors); with order=True also the input index of every rank."""
    torch = _torch()
    n = x.numel()
    for t, w in ((x, "x"), (y, "y"), (z, "z")):
        _check_dev(t, torch.float32, n, w)
    nc = peak.numel()
    _check_dev(peak, torch.int32, nc, "peak")
    _check_dev(key, torch.float32, nc, "key")
    keep = torch.empty(max(nc, 1), dtype=torch.int32, device=x.device)
    cnt = torch.zeros(1, dtype=torch.int32, device=x.device)
    od = torch.empty(max(nc, 1), dtype=torch.int32, device=x.device) if order else None
    lib = _lib.load()
    _lib.check(lib.pgalf_dev_merging_peak(_dptr(x), _dptr(y), _dptr(z), n, _dptr(peak),
                                          _dptr(key), nc, float(link), _dptr(keep), _dptr(od),
                                          _dptr(cnt), _stream()))
    return (keep, cnt, od) if order else (keep, cnt)


def dev_status():
    """Synchronise torch's current stream and raise on kernel-detected errors."""
    lib = _lib.load()
    _lib.check(lib.pgalf_dev_status(_stream()))


SCHEDULE_KEYS = ("knn_packed", "knn_xcd_chunk", "tsc_unit_entries")
SMOOTH_DIRECT_MAX = 56  # default of the library


def set_schedule(bins=(), smooth_direct_max=SMOOTH_DIRECT_MAX):
    """Install a launch schedule (pgalf_set_schedule): ``bins`` is a list of
    dicts {nmax, knn_packed, knn_xcd_chunk, tsc_unit_entries} by ascending
    nmax; a workload of n particles takes the first bin with n <= nmax, the
    library defaults beyond the last.  An empty list restores the defaults.
    A batch keeps the schedule it was created under."""
    lib = _lib.load()
    arr = (_lib.pgalf_schedule_t * max(1, len(bins)))()
    for i, b in enumerate(bins):
        arr[i].nmax = int(b["nmax"])
        for key in SCHEDULE_KEYS:
            setattr(arr[i], key, int(b[key]))
    _lib.check(lib.pgalf_set_schedule(arr, len(bins), int(smooth_direct_max)))


def get_schedule(n):
    """(schedule dict, smooth_direct_max) that a workload of n particles gets."""
    lib = _lib.load()
    out = _lib.pgalf_schedule_t()
    dm = ctypes.c_int()
    _lib.check(lib.pgalf_get_schedule(int(n), ctypes.byref(out), ctypes.byref(dm)))
    d = {"nmax": out.nmax}
    d.update({key: getattr(out, key) for key in SCHEDULE_KEYS})
    return d, dm.value


def timing(enable=None, reset=False):
    lib = _lib.load()
    if enable is not None:
        lib.pgalf_timing_enable(1 if enable else 0)
    if reset:
        lib.pgalf_timing_reset()


def timing_get(phase):
    lib = _lib.load()
    ms = ctypes.c_double()
    cnt = ctypes.c_int()
    lib.pgalf_timing_get(phase.encode(), ctypes.byref(ms), ctypes.byref(cnt))
    return ms.value, cnt.value


__all__ = ["PgalfError", "PARTICLE_DTYPE", "make_particles", "padded_shape",
           "assign_density_TSC", "gaussian_Smoothing", "findDen", "findsphdensity",
           "starfindsphdensity", "near_lists", "lag_density_stage", "grid_struct",
           "dev_tsc", "dev_smooth", "dev_findden", "dev_knn", "dev_density_stage",
           "dev_density_stage_backward", "dev_peak_climb", "dev_status", "Batch", "timing",
           "timing_get", "CORE_DTYPE", "grid_cores", "finddenpeak", "dev_grid_cores",
           "dev_finddenpeak", "set_schedule", "get_schedule", "MergingPeak",
           "dev_merging_peak"]


def dev_find_core_density(den, nbr, peak, types, mass):
    """FindCoreDensity on device (pgalf_dev_find_core_density).  den float32
    [n], nbr int64 [n, k], peak int32 [numcore], types int32 [n], mass
    float64 [n].  Returns a dict of device tensors: coredensity, nummem,
    numstar, starmass (per core), haloid, flag (per particle)."""
    torch = _torch()
    n = den.numel()
    k = nbr.shape[1]
    nc = peak.numel()
    _check_dev(den, torch.float32, n, "den")
    _check_dev(nbr, torch.int64, n * k, "nbr")
    _check_dev(peak, torch.int32, nc, "peak")
    _check_dev(types, torch.int32, n, "types")
    _check_dev(mass, torch.float64, n, "mass")
    dev = den.device
    o = dict(coredensity=torch.empty(max(nc, 1), dtype=torch.float32, device=dev),
             nummem=torch.empty(max(nc, 1), dtype=torch.int32, device=dev),
             numstar=torch.empty(max(nc, 1), dtype=torch.int32, device=dev),
             starmass=torch.empty(max(nc, 1), dtype=torch.float32, device=dev),
             haloid=torch.empty(max(n, 1), dtype=torch.int32, device=dev),
             flag=torch.empty(max(n, 1), dtype=torch.uint8, device=dev))
    lib = _lib.load()
    rc = lib.pgalf_dev_find_core_density(_dptr(den), n, _dptr(nbr), int(k), _dptr(types),
                                         _dptr(mass), _dptr(peak), nc, _dptr(o["coredensity"]),
                                         _dptr(o["nummem"]), _dptr(o["numstar"]),
                                         _dptr(o["starmass"]), _dptr(o["haloid"]),
                                         _dptr(o["flag"]), _stream())
    _lib.check(rc)
    for key in ("coredensity", "nummem", "numstar", "starmass"):
        o[key] = o[key][:nc]
    for key in ("haloid", "flag"):
        o[key] = o[key][:n]
    return o


def dev_shell_thresholds(den, coredensity):
    """The shell thresholds from device densities (pgalf_dev_shell_thresholds);
    a float32 numpy array."""
    torch = _torch()
    n, nc = den.numel(), coredensity.numel()
    _check_dev(den, torch.float32, n, "den")
    _check_dev(coredensity, torch.float32, nc, "coredensity")
    thr = np.zeros(MAX_SHELL_DIVIDE, dtype=np.float32)
    cnt = ctypes.c_int(0)
    lib = _lib.load()
    _lib.check(lib.pgalf_dev_shell_thresholds(_dptr(den), n, _dptr(coredensity), nc, _ptr(thr),
                                              ctypes.byref(cnt), _stream()))
    return thr[:cnt.value].copy()


def dev_shell_particles_peaks(den, nbr, peak, coredensity, thr, flag, haloid, maxcores=None):
    """Shells on device (pgalf_dev_shell_particles_peaks): flag and haloid
    (device tensors from dev_find_core_density) are updated in place.
    Returns (nshell, part_ptr, part, core_ptr, cores) as device tensors."""
    torch = _torch()
    n = den.numel()
    k = nbr.shape[1]
    nc = peak.numel()
    _check_dev(den, torch.float32, n, "den")
    _check_dev(nbr, torch.int64, n * k, "nbr")
    _check_dev(peak, torch.int32, nc, "peak")
    _check_dev(coredensity, torch.float32, nc, "coredensity")
    _check_dev(flag, torch.uint8, n, "flag")
    _check_dev(haloid, torch.int32, n, "haloid")
    thr = np.ascontiguousarray(thr, dtype=np.float32)
    dev = den.device
    maxshell = nc * max(len(thr), 1) + 1
    # room for a core listed by every sweep that reaches it (the 4M-particle
    # halo's 38,248 cores need 5.8M entries over 15 thresholds); more on
    # demand below
    maxcores = 16 * nc * (len(thr) + 1) + 16 if maxcores is None else maxcores
    pp = torch.zeros(maxshell + 1, dtype=torch.int64, device=dev)
    cp = torch.zeros(maxshell + 1, dtype=torch.int64, device=dev)
    part = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
    flag0, haloid0 = flag.clone(), haloid.clone()
    lib = _lib.load()
    while True:
        cores = torch.zeros(max(maxcores, 1), dtype=torch.int32, device=dev)
        sh = pgalf_shells_t(_dptr(pp), _dptr(part), _dptr(cp), _dptr(cores), maxshell, maxcores,
                            0)
        rc = lib.pgalf_dev_shell_particles_peaks(_dptr(den), n, _dptr(nbr), int(k), _dptr(peak),
                                                 _dptr(coredensity), nc, _ptr(thr), len(thr),
                                                 _dptr(flag), _dptr(haloid), ctypes.byref(sh),
                                                 _stream())
        # a core listed by several shells of one threshold can outgrow the
        # core entries: start again from the input state with more room
        if rc == 6 and maxcores < 1024 * (nc + 1) * (len(thr) + 1):
            flag.copy_(flag0)
            haloid.copy_(haloid0)
            maxcores *= 4
            continue
        _lib.check(rc)
        return sh.nshell, pp, part, cp, cores


def dev_catalog_watershed(off, core_off, den, nbr, peak, types, mass, small_max=0):
    """FindCoreDensity + thresholds + shells of every halo of a catalogue
    (pgalf_dev_catalog_watershed).  off / core_off: int64 numpy arrays
    (nhalo + 1); den float32, nbr int64 [n, k] with halo-local indices,
    peak int32 (halo-local), types int32, mass float64: device tensors.
    Halos of at most small_max particles (0: the library's 131072) run one
    per workgroup, the rest through the one-halo path on four host threads.
    Returns a dict of device tensors (coredensity, nummem, numstar, starmass,
    haloid, flag, part_ptr, part, core_ptr, cores) and host arrays
    (shell_base, core_base, nshell, status)."""
    torch = _torch()
    off = np.ascontiguousarray(off, dtype=np.int64)
    core_off = np.ascontiguousarray(core_off, dtype=np.int64)
    nhalo = len(off) - 1
    n, nc = int(off[-1]), int(core_off[-1])
    k = nbr.shape[1]
    _check_dev(den, torch.float32, n, "den")
    _check_dev(nbr, torch.int64, n * k, "nbr")
    _check_dev(peak, torch.int32, nc, "peak")
    _check_dev(types, torch.int32, n, "types")
    _check_dev(mass, torch.float64, n, "mass")
    lib = _lib.load()
    sbase = np.zeros(max(nhalo, 1), np.int64)
    cbase = np.zeros(max(nhalo, 1), np.int64)
    ccap = np.zeros(max(nhalo, 1), np.int64)
    tot_c = ctypes.c_longlong(0)
    tot_s = lib.pgalf_catalog_shell_layout(nhalo, _ptr(off), _ptr(core_off), _ptr(sbase),
                                           _ptr(cbase), _ptr(ccap), ctypes.byref(tot_c))
    if tot_s < 0:
        _lib.check()
    dev = den.device
    e = lambda m, dt: torch.empty(max(m, 1), dtype=dt, device=dev)  # noqa: E731
    o = dict(coredensity=e(nc, torch.float32), nummem=e(nc, torch.int32),
             numstar=e(nc, torch.int32), starmass=e(nc, torch.float32),
             haloid=e(n, torch.int32), flag=e(n, torch.uint8),
             part_ptr=e(tot_s, torch.int64), part=e(n, torch.int32),
             core_ptr=e(tot_s, torch.int64), cores=e(tot_c.value, torch.int32))
    nshell = np.zeros(max(nhalo, 1), np.int64)
    status = np.zeros(max(nhalo, 1), np.int32)
    rc = lib.pgalf_dev_catalog_watershed(
        nhalo, _ptr(off), _ptr(core_off), _dptr(den), _dptr(nbr), int(k), _dptr(types),
        _dptr(mass), _dptr(peak), _dptr(o["coredensity"]), _dptr(o["nummem"]),
        _dptr(o["numstar"]), _dptr(o["starmass"]), _dptr(o["haloid"]), _dptr(o["flag"]),
        _dptr(o["part_ptr"]), _dptr(o["part"]), _dptr(o["core_ptr"]), _dptr(o["cores"]),
        int(small_max), _ptr(nshell), _ptr(status), _stream())
    o.update(shell_base=sbase[:nhalo], core_base=cbase[:nhalo], nshell=nshell[:nhalo],
             status=status[:nhalo])
    _lib.check(rc)
    return o


def catalog_shells_of(o, off, h):
    """Halo h's shells from dev_catalog_watershed's result (host copies in
    o, numpy): a list of (particles, cores) arrays, local indices."""
    sb, cb, ns = int(o["shell_base"][h]), int(o["core_base"][h]), int(o["nshell"][h])
    pp = o["part_ptr"][sb:sb + ns + 1]
    cp = o["core_ptr"][sb:sb + ns + 1]
    part = o["part"][off[h]:off[h + 1]]
    cores = o["cores"][cb:]
    return [(part[pp[i]:pp[i + 1]], cores[cp[i]:cp[i + 1]]) for i in range(ns)]
