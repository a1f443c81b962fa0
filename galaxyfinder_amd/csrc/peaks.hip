// Integer peak stage on the GPU: the grid cores of lagFindStellarCore
// (NewGalFinder/nnost.c:1115-1271) and finddenpeak
// (NewGalFinder/subhaloden.mod6.c:416-492).  Both are exact integer /
// comparison work: the outputs are bit-identical to the reference's.
//
// Grid cores.  The reference chains particles per nearest grid node (head
// insertion in index order, so a chain runs from its largest index down),
// then for every interior node (nbuff/2 cells from each face) whose smoothed
// value exceeds the threshold and all 26 neighbours and whose chain is not
// empty, takes the first chain member of maximal densph -- the largest index
// among equal maxima -- unless the chain's last member, its smallest index,
// is particle 0 (`if(ibp > 0)`).  Cores come in (k, j, i) node order.  Here:
//   1. peak_flag_k   one thread per interior node: the strict 26-neighbour
//                    maximum above the threshold -> flag;
//   2. an exclusive scan of the flags gives the peak nodes in (k, j, i)
//      order (peak_list_k);
//   3. peak_member_k one thread per selected particle: its node (rint of the
//      FP64 offset in cells, as the reference), a binary search among the
//      peak nodes, then atomicMax of (ordered densph << 32 | index) -- the
//      maximal density, ties to the larger index -- and a flag when the
//      particle is particle 0;
//   4. core_compact_k one workgroup: the peaks whose chain is not empty and
//      does not hold particle 0, compacted in order.
// finddenpeak: one thread per particle (its density above the threshold, a
// star when jflag == 1, no neighbour strictly denser), then an exclusive scan
// compacts the cores in index order.
#include <cstdint>

#include "pgalf_internal.h"

namespace pgalf {
namespace {

inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// Order-preserving map of a float to uint32 (any non-NaN value maps above 0).
__device__ inline unsigned ord_f(float v) {
    const unsigned b = __float_as_uint(v);
    return (b >> 31) ? ~b : (b | 0x80000000u);
}

struct Interior {
    int i0, j0, k0, ni, nj, nk;  // first node and extent on each axis
    long long mx, ny;            // padded row and rows per plane
    __host__ __device__ long long count() const { return (long long)ni * nj * nk; }
    // padded flat index of interior node t (t in (k, j, i) order)
    __host__ __device__ long long cell(long long t) const {
        const long long i = t % ni, j = (t / ni) % nj, k = t / ((long long)ni * nj);
        return (i0 + i) + mx * ((j0 + j) + ny * (k0 + k));
    }
};

__global__ void peak_flag_k(const float *den, Interior in, double threshold, unsigned *flag) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= in.count()) return;
    const long long c = in.cell(t);
    const float v = den[c];
    bool peak = (double)v > threshold;
#pragma unroll
    for (int dk = -1; dk <= 1; dk++)
#pragma unroll
        for (int dj = -1; dj <= 1; dj++)
#pragma unroll
            for (int di = -1; di <= 1; di++) {
                if (di == 0 && dj == 0 && dk == 0) continue;
                peak &= v > den[c + di + in.mx * (dj + in.ny * dk)];
            }
    flag[t] = peak ? 1u : 0u;
}

// Peak nodes in (k, j, i) order; their per-peak state cleared.
__global__ void peak_list_k(const unsigned *flag, const unsigned *pos, Interior in,
                            long long *peak_cell, unsigned long long *best, unsigned *has0) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= in.count() || !flag[t]) return;
    const unsigned p = pos[t];
    peak_cell[p] = in.cell(t);
    best[p] = 0ull;
    has0[p] = 0u;
}

template <class P>
__global__ void peak_member_k(const P *x, const P *y, const P *z, const int *type, long long np,
                              int ptype, const float *densph, Grid g, const long long *peak_cell,
                              const unsigned *npeak, unsigned long long *best, unsigned *has0) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    if (ptype == PGALF_TYPE_STAR) {
        if (type[i] != PGALF_TYPE_STAR) return;
    } else if (ptype != PGALF_TYPE_ALL) {
        return;  // the reference chains stars or all particles only
    }
    const long long a = (long long)rint(((double)x[i] - g.x0) / g.h);
    const long long b = (long long)rint(((double)y[i] - g.y0) / g.h);
    const long long c = (long long)rint(((double)z[i] - g.z0) / g.h);
    if (a < 0 || a >= g.nx || b < 0 || b >= g.ny || c < 0 || c >= g.nz) return;
    const long long cell = a + g.mx() * (b + (long long)g.ny * c);
    // lower bound among the sorted peak nodes
    long long lo = 0, hi = (long long)*npeak;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (peak_cell[mid] < cell) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= (long long)*npeak || peak_cell[lo] != cell) return;
    atomicMax(&best[lo], ((unsigned long long)ord_f(densph[i]) << 32) | (unsigned long long)i);
    if (i == 0) has0[lo] = 1u;
}

// One workgroup of 1024: the peaks with a non-empty chain not holding
// particle 0, in order, to core[0 .. *ncore) (at most maxcore written).
__global__ void __launch_bounds__(1024) core_compact_k(const unsigned *npeak,
                                                       const unsigned long long *best,
                                                       const unsigned *has0, int *core,
                                                       long long maxcore, unsigned *ncore) {
    __shared__ unsigned s_wave[16];
    __shared__ unsigned s_base;
    const unsigned n = *npeak;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_base = 0;
    __syncthreads();
    for (unsigned c0 = 0; c0 < n; c0 += 1024) {
        const unsigned p = c0 + threadIdx.x;
        const bool keep = p < n && best[p] != 0ull && !has0[p];
        const unsigned long long m = __ballot(keep);
        if (lane == 0) s_wave[wave] = (unsigned)__popcll(m);
        __syncthreads();
        unsigned before = s_base;
        for (int w = 0; w < wave; w++) before += s_wave[w];
        const unsigned rank = before + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
        if (keep && rank < maxcore) core[rank] = (int)(unsigned)(best[p] & 0xffffffffull);
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned tot = 0;
            for (int w = 0; w < 16; w++) tot += s_wave[w];
            s_base += tot;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *ncore = s_base;
}

__global__ void denpeak_flag_k(const float *den, int k, const long long *nbr, long long np,
                               const int *type, int jflag, double threshold, unsigned *flag) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float v = den[i];
    bool peak = (double)v > threshold && (jflag != 1 || type[i] == PGALF_TYPE_STAR);
    if (peak) {
        const long long *li = nbr + i * k;
        for (int t = 0; t < k; t++) peak &= !(den[li[t]] > v);
    }
    flag[i] = peak ? 1u : 0u;
}

__global__ void denpeak_scatter_k(const unsigned *flag, const unsigned *pos, long long np,
                                  int *core, long long maxcore) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np || !flag[i]) return;
    if (pos[i] < maxcore) core[pos[i]] = (int)i;
}

}  // namespace

template <class P>
hipError_t grid_cores(const P *x, const P *y, const P *z, const int *type, long long np,
                      int ptype, const float *densph, const Grid &g, const float *den,
                      double threshold, int nbuff, int *core, long long maxcore,
                      unsigned *ncore, Workspace &ws, hipStream_t s) {
    WsScope scope(ws);
    const int h = nbuff / 2;
    Interior in;
    in.i0 = in.j0 = in.k0 = h;
    in.ni = g.nx - 2 * h;
    in.nj = g.ny - 2 * h;
    in.nk = g.nz - 2 * h;
    in.mx = g.mx();
    in.ny = g.ny;
    // an interior node's 26 neighbours must lie inside the grid
    if (h < 1 || in.ni < 1 || in.nj < 1 || in.nk < 1) return fill_u32(ncore, 1, 0u, s);
    const long long nint = in.count();
    unsigned *flag = ws.alloc_n<unsigned>(nint), *pos = ws.alloc_n<unsigned>(nint);
    unsigned *npeak = ws.alloc_n<unsigned>(1);
    const long long cap = nint;  // peak nodes: at most one per interior node
    long long *peak_cell = ws.alloc_n<long long>(cap);
    unsigned long long *best = ws.alloc_n<unsigned long long>(cap);
    unsigned *has0 = ws.alloc_n<unsigned>(cap);
    if (!flag || !pos || !npeak || !peak_cell || !best || !has0) return hipErrorOutOfMemory;
    peak_flag_k<<<nblk(nint, 256), 256, 0, s>>>(den, in, threshold, flag);
    PGALF_HIP_TRY(exclusive_scan_u32(flag, pos, nint, npeak, ws, s));
    peak_list_k<<<nblk(nint, 256), 256, 0, s>>>(flag, pos, in, peak_cell, best, has0);
    if (np > 0)
        peak_member_k<P><<<nblk(np, 256), 256, 0, s>>>(x, y, z, type, np, ptype, densph, g,
                                                       peak_cell, npeak, best, has0);
    core_compact_k<<<1, 1024, 0, s>>>(npeak, best, has0, core, maxcore, ncore);
    return hipGetLastError();
}

hipError_t find_den_peak(const float *den, int k, const long long *nbr, long long np,
                         const int *type, int jflag, double threshold, int *core,
                         long long maxcore, unsigned *ncore, Workspace &ws, hipStream_t s) {
    if (np <= 0) return fill_u32(ncore, 1, 0u, s);
    WsScope scope(ws);
    unsigned *flag = ws.alloc_n<unsigned>(np), *pos = ws.alloc_n<unsigned>(np);
    if (!flag || !pos) return hipErrorOutOfMemory;
    denpeak_flag_k<<<nblk(np, 256), 256, 0, s>>>(den, k, nbr, np, type, jflag, threshold, flag);
    PGALF_HIP_TRY(exclusive_scan_u32(flag, pos, np, ncore, ws, s));
    denpeak_scatter_k<<<nblk(np, 256), 256, 0, s>>>(flag, pos, np, core, maxcore);
    return hipGetLastError();
}

template hipError_t grid_cores<float>(const float *, const float *, const float *, const int *,
                                      long long, int, const float *, const Grid &,
                                      const float *, double, int, int *, long long, unsigned *,
                                      Workspace &, hipStream_t);
template hipError_t grid_cores<double>(const double *, const double *, const double *,
                                       const int *, long long, int, const float *, const Grid &,
                                       const float *, double, int, int *, long long, unsigned *,
                                       Workspace &, hipStream_t);

}  // namespace pgalf
