"""The library called from C exactly as NewGalFinder would call it: the
harness tests/c_harness/pgalf_harness.c includes include/pgalf_density.h,
links -lpgalf_density and calls assign_density_TSC, findDen,
gaussian_Smoothing, findsphdensity and starfindsphdensity with host arrays
of 88-byte particle records.  Its outputs on the reference fixtures
(tests/golden/tsc_small.npz, knn_small.npz) are checked against the
reference's outputs and the oracle:
  * deposit: normwise <= 1e-6, pointwise <= 1e-5 against the reference grid;
  * findDen: bit-exact against the reference;
  * smoothing: normwise <= 1e-6 against the FP64 oracle;
  * findsphdensity lists bit-exact against the reference's cold-start
    Find_Near_PreCopied lists, densities bit-exact against the oracle;
  * starfindsphdensity densities and lists bit-exact against the oracle.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from galaxyfinder_amd import density as D  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden")
HARNESS = os.path.join(ROOT, "tests", "c_harness", "pgalf_harness")


def test_c_caller_on_reference_fixtures(tmp_path):
    if not os.path.exists(HARNESS):
        pytest.fail("C harness not built (make -C tests/c_harness, run by build())")
    f = np.load(os.path.join(GOLD, "tsc_small.npz"))
    kf = np.load(os.path.join(GOLD, "knn_small.npz"))
    nx, ny, nz, x0, y0, z0, h = f["grid"]
    nx, ny, nz = int(nx), int(ny), int(nz)
    pos, mass, types = f["pos"], f["mass"], f["types"]
    n, k, RG = len(pos), 32, 2.0 * h
    assert np.array_equal(kf["pos"], pos)
    D.make_particles(pos, mass, types).tofile(tmp_path / "particles.bin")
    np.ascontiguousarray(f["den_p5_t1"], dtype=np.float32).tofile(tmp_path / "grid_in.bin")
    (tmp_path / "params.txt").write_text("%d %d %d %d %.17g %.17g %.17g %.17g %.9g %d %d\n"
                                         % (n, nx, ny, nz, x0, y0, z0, h, RG, k, 5))
    print("[case] C harness, N=%d grid=%dx%dx%d" % (n, nx, ny, nz), flush=True)
    r = subprocess.run([HARNESS, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    shape = D.padded_shape(nx, ny, nz)
    rd = lambda name, dt: np.fromfile(tmp_path / name, dtype=dt)  # noqa: E731
    # deposit
    got = rd("tsc.bin", np.float32).reshape(shape)[..., :nx].astype(np.float64)
    ref = f["den_p5_t1"][..., :nx].astype(np.float64)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-6
    mask = np.abs(ref) >= 1e-3 * np.abs(ref).max()
    assert (np.abs(got - ref)[mask] / np.abs(ref)[mask]).max() <= 1e-5
    # read-back of the reference grid
    assert np.array_equal(rd("findden.bin", np.float32).view(np.int32),
                          f["densph_p5"].view(np.int32))
    # smoothing of the reference grid
    sm = rd("smooth.bin", np.float32).reshape(shape)[..., :nx].astype(np.float64)
    ref = O.gsmooth_f64(np.ascontiguousarray(f["den_p5_t1"][..., :nx], dtype=np.float64), h, RG)
    assert np.linalg.norm(sm - ref) / np.linalg.norm(ref) <= 1e-6
    # kNN densities and lists
    near = rd("near.bin", np.int64).reshape(n, k)
    assert np.array_equal(near, kf["idx_cold"])
    oidx, od = O.findsphdensity(pos, k)
    assert np.array_equal(rd("sphden.bin", np.float32).view(np.int32), od.view(np.int32))
    sidx, sden = O.starfindsphdensity(pos, mass, types, k)
    assert np.array_equal(rd("star_near.bin", np.int64).reshape(n, k), sidx)
    assert np.array_equal(rd("star_den.bin", np.float32).view(np.int32), sden.view(np.int32))
