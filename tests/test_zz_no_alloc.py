"""pgalf_dev_knn_noalloc (static scratch, no allocation) against
pgalf_dev_knn, through the C-ABI: bit-identical lists and W4 densities on
two calls in a row, the second smaller than the first (the scratch keeps the
first call's contents), unweighted and mass-weighted, all candidates and a
star-only candidate subset; sizes above the scratch's limit are rejected
before any device work."""
import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from galaxyfinder_amd import _lib, halos  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print("[case] " + msg, flush=True)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _knn(fn, x, y, z, m, types, ptype, k, mode):
    import torch
    n = x.numel()
    nbr = torch.full((n, k), -7, dtype=torch.int64, device="cuda")
    rho = torch.full((n,), -7.0, dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(_p(x), _p(y), _p(z), _p(m), _p(types), n, ptype, k, _p(nbr), _p(rho), mode, stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    return nbr.cpu().numpy(), rho.cpu().numpy()


def test_noalloc_knn_equals_dev_knn_on_two_calls():
    import torch
    lib = _lib.load()
    assert lib.pgalf_dev_knn_noalloc_max() == 524288
    cases = ((400000, 32, 1, False, 61), (70000, 17, 2, True, 62))
    for n, k, mode, subset, seed in cases:
        say("noalloc kNN n=%d k=%d rho_mode=%d star subset=%s" % (n, k, mode, subset))
        pos, mass = halos.small_halo(n, seed=seed)
        mass = (mass * np.random.default_rng(seed).uniform(0.5, 1.5, n)).astype(np.float32)
        x, y, z = (torch.from_numpy(np.ascontiguousarray(pos[:, a])).cuda() for a in range(3))
        m = torch.from_numpy(mass).cuda()
        types, ptype = None, 5
        if subset:
            types = torch.from_numpy((np.arange(n) % 3 == 0).astype(np.int32) + 1).cuda()
            ptype = 1
        a = _knn(lib.pgalf_dev_knn, x, y, z, m, types, ptype, k, mode)
        b = _knn(lib.pgalf_dev_knn_noalloc, x, y, z, m, types, ptype, k, mode)
        assert np.array_equal(a[0], b[0]), "neighbour lists differ"
        assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)), "densities differ"
        assert (b[0] >= 0).all()


def test_noalloc_knn_rejects_sizes_beyond_its_scratch():
    import torch
    lib = _lib.load()
    n = int(lib.pgalf_dev_knn_noalloc_max()) + 1
    x = torch.zeros(n, dtype=torch.float32, device="cuda")
    rc = lib.pgalf_dev_knn_noalloc(_p(x), _p(x), _p(x), None, None, n, 5, 32, None, None, 0,
                                   None)
    assert rc == 1  # PGALF_ERR_ARG
    assert b"static scratch" in lib.pgalf_last_error_string()
