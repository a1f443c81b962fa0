// Integer peak stage on the GPU: the grid cores of lagFindStellarCore
// (NewGalFinder/nnost.c:1115-1271), finddenpeak
// (NewGalFinder/subhaloden.mod6.c:416-492) and MergingPeak (:341-415).  All
// are exact integer / comparison work: the outputs are bit-identical to the
// reference's.
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
//   2. an exclusive scan of the flags ranks the peak nodes in (k, j, i)
//      order; peak_list_k clears their state;
//   3. peak_member_k one thread per selected particle: its node (rint of the
//      FP64 offset in cells, as the reference), the node's flag and rank
//      read directly, then atomicMax of (ordered densph << 32 | index) -- the
//      maximal density, ties to the larger index -- and a flag when the
//      particle is particle 0;
//   4. core_compact_k one workgroup: the peaks whose chain is not empty and
//      does not hold particle 0, compacted in order.
// finddenpeak: one thread per particle (its density above the threshold, a
// star when jflag == 1, no neighbour strictly denser; lists of k <= 32 staged
// in LDS by coalesced loads), then an exclusive scan compacts the cores in
// index order.
#include <cfloat>
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

// Per-peak state cleared (peak p is the p-th flagged node in (k, j, i) order).
__global__ void peak_list_k(const unsigned *flag, const unsigned *pos, Interior in,
                            unsigned long long *best, unsigned *has0) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= in.count() || !flag[t]) return;
    const unsigned p = pos[t];
    best[p] = 0ull;
    has0[p] = 0u;
}

// The particle's node (rint of the FP64 offset in cells, as the reference);
// only interior nodes can be peaks, and their flags and scan positions give
// the peak's rank directly.
template <class P>
__global__ void peak_member_k(const P *x, const P *y, const P *z, const int *type, long long np,
                              int ptype, const float *densph, Grid g, Interior in,
                              const unsigned *flag, const unsigned *pos,
                              unsigned long long *best, unsigned *has0) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    if (ptype == PGALF_TYPE_STAR) {
        if (type[i] != PGALF_TYPE_STAR) return;
    } else if (ptype != PGALF_TYPE_ALL) {
        return;  // the reference chains stars or all particles only
    }
    const long long a = (long long)rint(((double)x[i] - g.x0) / g.h) - in.i0;
    const long long b = (long long)rint(((double)y[i] - g.y0) / g.h) - in.j0;
    const long long c = (long long)rint(((double)z[i] - g.z0) / g.h) - in.k0;
    if (a < 0 || a >= in.ni || b < 0 || b >= in.nj || c < 0 || c >= in.nk) return;
    const long long t = a + (long long)in.ni * (b + (long long)in.nj * c);
    if (!flag[t]) return;
    const unsigned p = pos[t];
    atomicMax(&best[p], ((unsigned long long)ord_f(densph[i]) << 32) | (unsigned long long)i);
    if (i == 0) has0[p] = 1u;
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

// k <= 32: a block of 256 particles stages its 256 lists (low 32 bits of
// each index; particle indices are below 2^31) in LDS with coalesced loads,
// two list rows per wave instruction, rows padded to 33 entries.
constexpr int kDenPeakRows = 256;
__global__ void __launch_bounds__(kDenPeakRows)
denpeak_flag_lds_k(const float *den, int k, const long long *nbr, long long np, const int *type,
                   int jflag, double threshold, unsigned *flag) {
    __shared__ int s_idx[kDenPeakRows * 33];
    const long long base = (long long)blockIdx.x * kDenPeakRows;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane & 31;
    for (int r = wave * 2 + (lane >> 5); r < kDenPeakRows; r += 2 * (kDenPeakRows / 64)) {
        if (t < k && base + r < np) s_idx[r * 33 + t] = (int)nbr[(base + r) * k + t];
    }
    __syncthreads();
    const long long i = base + threadIdx.x;
    if (i >= np) return;
    const float v = den[i];
    bool peak = (double)v > threshold && (jflag != 1 || type[i] == PGALF_TYPE_STAR);
    if (peak) {
        const int *li = s_idx + threadIdx.x * 33;
        for (int u = 0; u < k; u++) peak &= !(den[li[u]] > v);
    }
    flag[i] = peak ? 1u : 0u;
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

// ---- MergingPeak (subhaloden.mod6.c:341-415) ---------------------------------
// Cores are ranked by key, largest first, equal keys in input order (a
// stable radix sort of the inverted order-preserving key bits; -0 counts as
// +0, as the reference's comparison does).  The reference then walks its
// friends-of-friends tree from each unclaimed core in rank order, claiming
// the whole group: the kept cores are the first-ranked core of each
// connected component of the link graph, in rank order.  Here the
// components come from a lock-free union-find whose roots are the smallest
// rank of their component (a larger root is always hooked under a smaller
// one), over the pairs found in a uniform cell grid: cells at least the link
// length wide (the 27-cell neighbourhood then holds every pair within the
// link), at most 1024 per axis, keys sorted by a stable radix sort.
constexpr int kMergeCellBits = 10;
constexpr int kMergeCells = 1 << kMergeCellBits;

__global__ void merge_key_k(const float *key, int n, unsigned *skey, unsigned *sval) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = key[i] == 0.0f ? 0.0f : key[i];
    skey[i] = ~ord_f(v);
    sval[i] = (unsigned)i;
}

// Positions in rank order; a peak index outside [0, np) raises
// DEV_ERR_BAD_INDEX and gives NaN coordinates, which link to nothing.
template <class P>
__global__ void merge_gather_k(const P *x, const P *y, const P *z, const int *peak,
                               long long np, const unsigned *order, int n, double *rx,
                               double *ry, double *rz, unsigned *status) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int i = (int)order[r];
    const long long p = peak ? (long long)peak[i] : (long long)i;
    if (p < 0 || p >= np) {
        atomicOr(status, (unsigned)DEV_ERR_BAD_INDEX);
        rx[r] = ry[r] = rz[r] = __longlong_as_double(0x7ff8000000000000ll);
        return;
    }
    rx[r] = (double)x[p];
    ry[r] = (double)y[p];
    rz[r] = (double)z[p];
}

// One workgroup: the bounding box of the ranked positions, then the cell
// width (bb[6]) and origin (bb[0..2]).
__global__ void __launch_bounds__(1024) merge_bbox_k(const double *rx, const double *ry,
                                                     const double *rz, int n, double link,
                                                     double *bb) {
    __shared__ double s_lo[3][16], s_hi[3][16];
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        const double v[3] = {rx[r], ry[r], rz[r]};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = fmin(lo[a], v[a]);
            hi[a] = fmax(hi[a], v[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        for (int o = 32; o > 0; o >>= 1) {
            lo[a] = fmin(lo[a], __shfl_xor(lo[a], o, 64));
            hi[a] = fmax(hi[a], __shfl_xor(hi[a], o, 64));
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int a = 0; a < 3; a++) {
            s_lo[a][wave] = lo[a];
            s_hi[a][wave] = hi[a];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) / 64;
        double ext = 0.0;
        for (int a = 0; a < 3; a++) {
            double l = s_lo[a][0], h = s_hi[a][0];
            for (int w = 1; w < nw; w++) {
                l = fmin(l, s_lo[a][w]);
                h = fmax(h, s_hi[a][w]);
            }
            bb[a] = l;
            ext = fmax(ext, h - l);
        }
        // wider than the link by a margin that covers the rounding of the
        // link test and of the cell coordinates
        bb[6] = fmax(link * (1.0 + 1.0 / 1048576.0), ext / (double)(kMergeCells - 8));
    }
}

__device__ inline int merge_cell(double v, double lo, double w) {
    const double c = floor((v - lo) / w);
    if (!(c >= 0.0)) return 0;  // also NaN
    return c > (double)(kMergeCells - 1) ? kMergeCells - 1 : (int)c;
}

__global__ void merge_cell_k(const double *rx, const double *ry, const double *rz, int n,
                             const double *bb, unsigned *ckey, unsigned *cval) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double w = bb[6];
    const unsigned cx = (unsigned)merge_cell(rx[r], bb[0], w);
    const unsigned cy = (unsigned)merge_cell(ry[r], bb[1], w);
    const unsigned cz = (unsigned)merge_cell(rz[r], bb[2], w);
    ckey[r] = cx | (cy << kMergeCellBits) | (cz << (2 * kMergeCellBits));
    cval[r] = (unsigned)r;
}

__global__ void merge_init_k(int *parent, int n) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) parent[r] = r;
}

__device__ inline int uf_find(int *parent, int x) {
    for (;;) {
        const int p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p == x) return x;
        x = p;
    }
}

// Joins the sets of a and b; parents only ever decrease, so every retry
// after a lost race starts from smaller roots and the loop ends.
__device__ inline void uf_union(int *parent, int a, int b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&parent[b], b, a) == b) return;
    }
}

// The link test of ost.c:328-334: FP64 squared distance, its float square
// root, compared in FP64 with the link length.
__device__ inline bool fof_linked(double ax, double ay, double az, double bx, double by,
                                  double bz, double link) {
    const double dx = ax - bx, dy = ay - by, dz = az - bz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    return (double)sqrtf((float)d2) <= link;
}

__global__ void merge_link_k(const double *rx, const double *ry, const double *rz, int n,
                             const double *bb, const unsigned *cs, const unsigned *cr,
                             double link, int *parent) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double w = bb[6];
    const double px = rx[r], py = ry[r], pz = rz[r];
    const int cx = merge_cell(px, bb[0], w), cy = merge_cell(py, bb[1], w),
              cz = merge_cell(pz, bb[2], w);
    for (int dz = -1; dz <= 1; dz++) {
        const int z = cz + dz;
        if (z < 0 || z >= kMergeCells) continue;
        for (int dy = -1; dy <= 1; dy++) {
            const int y = cy + dy;
            if (y < 0 || y >= kMergeCells) continue;
            for (int dx = -1; dx <= 1; dx++) {
                const int x = cx + dx;
                if (x < 0 || x >= kMergeCells) continue;
                const unsigned key = (unsigned)x | ((unsigned)y << kMergeCellBits) |
                                     ((unsigned)z << (2 * kMergeCellBits));
                int lo = 0, hi = n;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (cs[mid] < key) lo = mid + 1;
                    else hi = mid;
                }
                for (int t = lo; t < n && cs[t] == key; t++) {
                    const int j = (int)cr[t];
                    if (j > r && fof_linked(px, py, pz, rx[j], ry[j], rz[j], link))
                        uf_union(parent, r, j);
                }
            }
        }
    }
}

__global__ void merge_root_k(const int *parent, int n, unsigned *flag) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) flag[r] = parent[r] == r ? 1u : 0u;
}

__global__ void merge_out_k(const unsigned *flag, const unsigned *pos, const unsigned *order,
                            int n, int *keep, int *order_out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    if (flag[r]) keep[pos[r]] = (int)order[r];
    if (order_out) order_out[r] = (int)order[r];
}

}  // namespace

template <class P>
hipError_t merging_peak(const P *x, const P *y, const P *z, long long np, const int *peak,
                        const float *key, int n, double link, int *keep, unsigned *nkeep,
                        int *order_out, Workspace &ws, hipStream_t s) {
    if (n <= 0) return fill_u32(nkeep, 1, 0u, s);
    WsScope scope(ws);
    unsigned *k0 = ws.alloc_n<unsigned>(n), *v0 = ws.alloc_n<unsigned>(n);
    unsigned *k1 = ws.alloc_n<unsigned>(n), *v1 = ws.alloc_n<unsigned>(n);
    double *rx = ws.alloc_n<double>(n), *ry = ws.alloc_n<double>(n), *rz = ws.alloc_n<double>(n);
    double *bb = ws.alloc_n<double>(8);
    unsigned *ck = ws.alloc_n<unsigned>(n), *cv = ws.alloc_n<unsigned>(n);
    int *parent = ws.alloc_n<int>(n);
    if (!k0 || !v0 || !k1 || !v1 || !rx || !ry || !rz || !bb || !ck || !cv || !parent)
        return hipErrorOutOfMemory;
    const unsigned nb = nblk(n, 256);
    merge_key_k<<<nb, 256, 0, s>>>(key, n, k0, v0);
    unsigned *ks = k0, *order = v0;
    PGALF_HIP_TRY(radix_sort_pairs(&ks, &order, k1, v1, n, 32, ws, s));
    // the other pair of buffers is free again: cell keys sort through it
    unsigned *kt = ks == k0 ? k1 : k0, *vt = order == v0 ? v1 : v0;
    merge_gather_k<P><<<nb, 256, 0, s>>>(x, y, z, peak, np, order, n, rx, ry, rz, ws.status());
    merge_bbox_k<<<1, 1024, 0, s>>>(rx, ry, rz, n, link, bb);
    merge_cell_k<<<nb, 256, 0, s>>>(rx, ry, rz, n, bb, ck, cv);
    unsigned *cs = ck, *cr = cv;
    PGALF_HIP_TRY(radix_sort_pairs(&cs, &cr, kt, vt, n, 3 * kMergeCellBits, ws, s));
    merge_init_k<<<nb, 256, 0, s>>>(parent, n);
    merge_link_k<<<nb, 256, 0, s>>>(rx, ry, rz, n, bb, cs, cr, link, parent);
    // flags and their positions reuse the cell-key buffer that is not live
    unsigned *flag = cs == ck ? kt : ck, *pos = cr == cv ? vt : cv;
    merge_root_k<<<nb, 256, 0, s>>>(parent, n, flag);
    PGALF_HIP_TRY(exclusive_scan_u32(flag, pos, n, nkeep, ws, s));
    merge_out_k<<<nb, 256, 0, s>>>(flag, pos, order, n, keep, order_out);
    return hipGetLastError();
}

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
    unsigned long long *best = ws.alloc_n<unsigned long long>(cap);
    unsigned *has0 = ws.alloc_n<unsigned>(cap);
    if (!flag || !pos || !npeak || !best || !has0) return hipErrorOutOfMemory;
    peak_flag_k<<<nblk(nint, 256), 256, 0, s>>>(den, in, threshold, flag);
    PGALF_HIP_TRY(exclusive_scan_u32(flag, pos, nint, npeak, ws, s));
    peak_list_k<<<nblk(nint, 256), 256, 0, s>>>(flag, pos, in, best, has0);
    if (np > 0)
        peak_member_k<P><<<nblk(np, 256), 256, 0, s>>>(x, y, z, type, np, ptype, densph, g, in,
                                                       flag, pos, best, has0);
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
    if (k <= 32 && np < (1LL << 31))
        denpeak_flag_lds_k<<<nblk(np, kDenPeakRows), kDenPeakRows, 0, s>>>(den, k, nbr, np, type,
                                                                          jflag, threshold, flag);
    else
        denpeak_flag_k<<<nblk(np, 256), 256, 0, s>>>(den, k, nbr, np, type, jflag, threshold,
                                                     flag);
    PGALF_HIP_TRY(exclusive_scan_u32(flag, pos, np, ncore, ws, s));
    denpeak_scatter_k<<<nblk(np, 256), 256, 0, s>>>(flag, pos, np, core, maxcore);
    return hipGetLastError();
}

template hipError_t merging_peak<float>(const float *, const float *, const float *,
                                        long long, const int *, const float *, int, double,
                                        int *, unsigned *, int *, Workspace &, hipStream_t);
template hipError_t merging_peak<double>(const double *, const double *, const double *,
                                         long long, const int *, const float *, int, double,
                                         int *, unsigned *, int *, Workspace &, hipStream_t);
template hipError_t grid_cores<float>(const float *, const float *, const float *, const int *,
                                      long long, int, const float *, const Grid &,
                                      const float *, double, int, int *, long long, unsigned *,
                                      Workspace &, hipStream_t);
template hipError_t grid_cores<double>(const double *, const double *, const double *,
                                       const int *, long long, int, const float *, const Grid &,
                                       const float *, double, int, int *, long long, unsigned *,
                                       Workspace &, hipStream_t);

}  // namespace pgalf
