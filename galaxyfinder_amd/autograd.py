"""Differentiable density stage: a ``torch.autograd.Function`` over the HIP
forward kernels (deposit -> smoothing -> read-back, kNN lists + W4 density)
and their hand-written adjoints (pgalf_dev_density_stage_backward).

    rho_tsc, rho_knn, nbr, den = DensityStage.apply(x, y, z, m, grid, RG, k, weighted)

x, y, z, m: float32 CUDA tensors (SoA).  rho_tsc is the read-back of the
smoothed TSC grid (findDen of gaussian_Smoothing of assign_density_TSC,
tsc_omp2.c / gsmooth.c); rho_knn the W4 kNN density over the k nearest
(nnost.c:536-555, mass-weighted as starfindsphdensity when ``weighted``,
unweighted as findsphdensity otherwise).  nbr (int64 [n, k]) and den (the
smoothed padded grid) are returned as non-differentiable outputs.

Gradients flow to x, y, z and m through both densities.  They are the
derivatives of the FP64 mathematical forward with the neighbour lists held
fixed (the lists are piecewise constant in the positions); the reference has
no adjoint, so this is an addition of this build.
"""
import torch

from . import density as D


class DensityStage(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, z, m, grid, RG, k, weighted=True):
        n = x.numel()
        dev = x.device
        x, y, z, m = (t.detach().contiguous() for t in (x, y, z, m))
        den = torch.empty(D.padded_shape(grid.nx, grid.ny, grid.nz), dtype=torch.float32,
                          device=dev)
        rho_t = torch.empty(n, dtype=torch.float32, device=dev)
        nbr = torch.empty((n, k), dtype=torch.int64, device=dev)
        rho_k = torch.empty(n, dtype=torch.float32, device=dev)
        D.dev_tsc(x, y, z, m, grid, den)
        D.dev_smooth(den, grid, RG)
        D.dev_findden(x, y, z, grid, den, rho_t)
        rho_mode = 2 if weighted else 1
        D.dev_knn(x, y, z, k, m=m, nbr=nbr, rho=rho_k, rho_mode=rho_mode)
        ctx.save_for_backward(x, y, z, m, den, nbr)
        ctx.grid, ctx.RG, ctx.k, ctx.rho_mode = grid, float(RG), int(k), rho_mode
        ctx.mark_non_differentiable(nbr, den)
        return rho_t, rho_k, nbr, den

    @staticmethod
    def backward(ctx, g_t, g_k, _g_nbr, _g_den):
        x, y, z, m, den, nbr = ctx.saved_tensors
        g_t = None if g_t is None else g_t.contiguous().float()
        g_k = None if g_k is None else g_k.contiguous().float()
        if g_t is None and g_k is None:
            return None, None, None, None, None, None, None, None
        gx, gy, gz, gm = D.dev_density_stage_backward(
            x, y, z, m, ctx.grid, ctx.RG, ctx.k, den, nbr, ctx.rho_mode, g_tsc=g_t, g_knn=g_k)
        return gx, gy, gz, gm, None, None, None, None


def density_stage(x, y, z, m, grid, RG, k=32, weighted=True):
    """Functional form of DensityStage.apply."""
    return DensityStage.apply(x, y, z, m, grid, RG, k, weighted)
