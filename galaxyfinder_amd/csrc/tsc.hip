// TSC mass deposit and TSC read-back for gfx950.
//
// Deposit (assign_density_TSC, NewGalFinder/tsc_omp2.c:166-506):
//   Tiles are TY x TZ whole padded rows of a grid (TY * TZ * mx <= kTileCap
//   nodes), so a particle's 27-node stencil touches at most 2 x 2 of them.
//   1. count:   every selected particle lists the tiles its stencil touches
//               -> per-tile counts (block-local LDS histograms, one global
//               add per block and tile);
//   2. scan:    tile offsets; tiles are cut into work units of <= `unit`
//               entries (the schedule's tsc_unit_entries) so a cusp tile is
//               shared by many workgroups;
//   3. scatter: the particles (position and mass) into per-tile bins
//               (ranges reserved per block), so the deposit reads its
//               entries coalesced;
//   4. deposit: one workgroup per unit accumulates its particles'
//               contributions that land inside the tile in an LDS int64
//               fixed-point tile (ds_add_u64).  Integer addition is
//               associative, so the grid is bit-identical run to run
//               whatever the order of the atomics.  A single-unit tile is
//               written straight into the grid: TZ runs of TY * mx
//               consecutive floats; multi-unit tiles are summed in a global
//               int64 side buffer;
//   5. rows:    the tiles the deposit did not write -- zeros for tiles
//               without entries, the side sums for multi-unit tiles.
//   The per-contribution value is the reference's FP32 product
//   (wx*wy)*wz with its float/double conversions (tsc_omp2.c:303-382); the
//   fixed-point scale 2^S is chosen from max|m| * count so no sum overflows.
//   Tiles span x in [0, mx), so pad columns are written as zero.
//
// Read-back (findDen, tsc_omp2.c:40-163): one thread per particle, 27
// gathers summed in the reference's order and precision -- bit-identical to
// the reference on the same grid (no FMA contraction: -ffp-contract=off).
#include "pgalf_internal.h"

namespace pgalf {
namespace {

// A deposit tile is TY x TZ whole padded rows (mx floats each) of one grid:
// its output is TZ runs of TY * mx consecutive floats of the grid, written
// straight from the workgroup's LDS accumulator, so no tile-shaped stores
// (partly written lines at every tile edge) and no second pass.  TY and TZ
// are powers of two with TY * TZ * mx <= kTileCap, as square as they come.
constexpr int kTileCap = 4352;       // nodes of one tile: 34 KB of int64 in LDS
constexpr int kLdsHistMax = 8192;    // tiles handled with LDS histograms
constexpr int kBlock = 256;
constexpr int kPerBlock = 8192;      // particles per count/scatter block

struct TileGeom {
    int mx, TY, TZ;  // row width, rows of a tile along y and z
    int nty, ntz;    // tiles along y and z
    __host__ __device__ long long ntiles() const { return (long long)nty * ntz; }
    __host__ __device__ int nodes() const { return mx * TY * TZ; }
};

struct Axis {
    long long nr;
    long long sg;
    float w1, w2, w3;
};

// Per-axis deposit weights (tsc_omp2.c:303-343).  with_mass: the x axis,
// whose weights carry the particle mass.
__device__ inline void deposit_axis(double p, double pm, bool with_mass, Axis &a) {
    double r = rint(p);
    a.nr = (long long)r;
    float off = (float)(p - (double)a.nr);
    a.sg = (long long)copysignf(1.0f, off);
    float d = (float)fabs((double)off);
    float d2 = d * d;
    if (with_mass) {
        double ph = pm * 0.5;
        a.w1 = (float)((0.75 - (double)d2) * pm);
        a.w3 = (float)(ph * (0.25 + (double)d * ((double)d - 1.0)));
        a.w2 = (float)((double)a.w3 + pm * (double)d);
    } else {
        a.w1 = (float)(0.75 - (double)d2);
        a.w3 = (float)(0.5 * (0.25 + (double)d * ((double)d - 1.0)));
        a.w2 = a.w3 + d;
    }
}

// Weight of node offset o (relative to the nearest node) along one axis.
__device__ inline float axis_weight(const Axis &a, long long o) {
    return o == 0 ? a.w1 : (o == a.sg ? a.w2 : a.w3);
}

template <class P>
__device__ inline bool particle_axes(const P *x, const P *y, const P *z, const P *m,
                                     long long i, const Grid &g, Axis ax[3]) {
    double pm = (double)m[i];
    deposit_axis(((double)x[i] - g.x0) / g.h, pm, true, ax[0]);
    deposit_axis(((double)y[i] - g.y0) / g.h, pm, false, ax[1]);
    deposit_axis(((double)z[i] - g.z0) / g.h, pm, false, ax[2]);
    // x nodes may reach the pad columns [nx, mx) like the reference's
    // unchecked writes do (memory of the padded row); y/z must stay inside.
    return ax[0].nr >= 1 && ax[0].nr <= g.mx() - 2 && ax[1].nr >= 1 &&
           ax[1].nr <= g.ny - 2 && ax[2].nr >= 1 && ax[2].nr <= g.nz - 2;
}

// Bin entry: the particle itself (position and mass), so that the deposit
// reads its entries with coalesced loads instead of gathering particles in
// input order.
template <class P>
struct alignas(4 * sizeof(P)) BinRec {
    P x, y, z, m;
};

template <class P>
__device__ inline void rec_axes(const BinRec<P> &r, const Grid &g, Axis ax[3]) {
    const double pm = (double)r.m;
    deposit_axis(((double)r.x - g.x0) / g.h, pm, true, ax[0]);
    deposit_axis(((double)r.y - g.y0) / g.h, pm, false, ax[1]);
    deposit_axis(((double)r.z - g.z0) / g.h, pm, false, ax[2]);
}

__device__ inline bool selected(const int *type, long long i, int ptype) {
    return ptype == PGALF_TYPE_ALL || type == nullptr || type[i] == ptype;
}

// Tiles of extent `ext` touched by nodes [nr-1, nr+1] along one axis.
__device__ inline void tile_span(long long nr, int ext, int &t0, int &t1) {
    t0 = (int)((nr - 1) / ext);
    t1 = (int)((nr + 1) / ext);
}

// Halo `h` of a batch, or the single halo `seg0` of a one-halo call.
__device__ inline HaloSeg seg_at(const HaloSeg *segs, const HaloSeg &seg0, int h) {
    return segs ? segs[h] : seg0;
}
// This block's particle chunk: from the batch's chunk list, or block
// `blockIdx.x` of `per` particles of the single halo.
__device__ inline PChunk chunk_at(const PChunk *chunks, const HaloSeg &seg0, int per) {
    if (chunks) return chunks[blockIdx.x];
    PChunk c;
    c.halo = 0;
    c.b0 = (long long)blockIdx.x * per;
    c.b1 = min(c.b0 + per, seg0.np);
    return c;
}

__host__ __device__ inline TileGeom tile_geom(const Grid &g) {
    TileGeom tg;
    tg.mx = (int)g.mx();
    const int rows = kTileCap / tg.mx;  // >= 1: checked by the host (tsc_grid_fits)
    int TY = 1, TZ = 1;
    while (2 * TY * TZ <= rows && TY * TZ < 256) {
        if (TZ <= TY) TZ *= 2;
        else TY *= 2;
    }
    tg.TY = TY;
    tg.TZ = TZ;
    tg.nty = (g.ny + TY - 1) / TY;
    tg.ntz = (g.nz + TZ - 1) / TZ;
    return tg;
}

// Per-block tile histogram of the chunk's halo (LDS when the halo has at
// most kLdsHistMax tiles), then global per-tile counts; per-halo selected
// count and max |m| for the fixed-point scale.
template <class P>
__global__ void __launch_bounds__(kBlock)
tsc_count_k(const P *x, const P *y, const P *z, const P *m, const int *type, int ptype,
            const HaloSeg *segs, HaloSeg seg0, const PChunk *chunks, unsigned *tcount,
            unsigned *nsel, unsigned long long *maxm_bits, unsigned *status, int use_lds) {
    extern __shared__ unsigned hist[];
    const PChunk c = chunk_at(chunks, seg0, kPerBlock);
    const HaloSeg sg = seg_at(segs, seg0, c.halo);
    const Grid g = sg.g;
    const TileGeom tg = tile_geom(g);
    const long long ntiles = tg.ntiles();
    if (use_lds) {
        for (long long t = threadIdx.x; t < ntiles; t += kBlock) hist[t] = 0;
        __syncthreads();
    }
    unsigned local_sel = 0;
    double local_max = 0.0;
    for (long long i = c.b0 + threadIdx.x; i < c.b1; i += kBlock) {
        if (!selected(type, i, ptype)) continue;
        Axis ax[3];
        if (!particle_axes(x, y, z, m, i, g, ax)) {
            atomicOr(status, (unsigned)DEV_ERR_OUT_OF_GRID);
            continue;
        }
        local_sel++;
        local_max = fmax(local_max, fabs((double)m[i]));
        int ty0, ty1, tz0, tz1;
        tile_span(ax[1].nr, tg.TY, ty0, ty1);
        tile_span(ax[2].nr, tg.TZ, tz0, tz1);
        for (int tz = tz0; tz <= tz1; tz++)
            for (int ty = ty0; ty <= ty1; ty++) {
                const long long t = ty + (long long)tg.nty * tz;
                if (use_lds) atomicAdd(&hist[t], 1u);
                else atomicAdd(&tcount[sg.toff + t], 1u);
            }
    }
    if (local_sel) atomicAdd(&nsel[c.halo], local_sel);
    if (local_max > 0.0)
        atomicMax(&maxm_bits[c.halo], (unsigned long long)__double_as_longlong(local_max));
    if (use_lds) {
        __syncthreads();
        for (long long t = threadIdx.x; t < ntiles; t += kBlock)
            if (hist[t]) atomicAdd(&tcount[sg.toff + t], hist[t]);
    }
}

template <class P>
__global__ void __launch_bounds__(kBlock)
tsc_scatter_k(const P *x, const P *y, const P *z, const P *m, const int *type, int ptype,
              const HaloSeg *segs, HaloSeg seg0, const PChunk *chunks, const unsigned *tstart,
              unsigned *tcursor, BinRec<P> *bins, int use_lds) {
    extern __shared__ unsigned sh[];
    const PChunk c = chunk_at(chunks, seg0, kPerBlock);
    const HaloSeg sg = seg_at(segs, seg0, c.halo);
    const Grid g = sg.g;
    const TileGeom tg = tile_geom(g);
    const long long ntiles = tg.ntiles();
    unsigned *hist = sh;
    unsigned *base = sh + (use_lds ? ntiles : 0);
    if (use_lds) {
        for (long long t = threadIdx.x; t < ntiles; t += kBlock) hist[t] = 0;
        __syncthreads();
        for (long long i = c.b0 + threadIdx.x; i < c.b1; i += kBlock) {
            if (!selected(type, i, ptype)) continue;
            Axis ax[3];
            if (!particle_axes(x, y, z, m, i, g, ax)) continue;
            int ty0, ty1, tz0, tz1;
            tile_span(ax[1].nr, tg.TY, ty0, ty1);
            tile_span(ax[2].nr, tg.TZ, tz0, tz1);
            for (int tz = tz0; tz <= tz1; tz++)
                for (int ty = ty0; ty <= ty1; ty++)
                    atomicAdd(&hist[ty + (long long)tg.nty * tz], 1u);
        }
        __syncthreads();
        for (long long t = threadIdx.x; t < ntiles; t += kBlock) {
            unsigned cnt = hist[t];
            base[t] = cnt ? tstart[sg.toff + t] + atomicAdd(&tcursor[sg.toff + t], cnt) : 0u;
            hist[t] = 0;
        }
        __syncthreads();
    }
    for (long long i = c.b0 + threadIdx.x; i < c.b1; i += kBlock) {
        if (!selected(type, i, ptype)) continue;
        Axis ax[3];
        if (!particle_axes(x, y, z, m, i, g, ax)) continue;
        BinRec<P> rec;
        rec.x = x[i];
        rec.y = y[i];
        rec.z = z[i];
        rec.m = m[i];
        int ty0, ty1, tz0, tz1;
        tile_span(ax[1].nr, tg.TY, ty0, ty1);
        tile_span(ax[2].nr, tg.TZ, tz0, tz1);
        for (int tz = tz0; tz <= tz1; tz++)
            for (int ty = ty0; ty <= ty1; ty++) {
                const long long t = ty + (long long)tg.nty * tz;
                unsigned slot;
                if (use_lds) slot = base[t] + atomicAdd(&hist[t], 1u);
                else slot = tstart[sg.toff + t] + atomicAdd(&tcursor[sg.toff + t], 1u);
                bins[slot] = rec;
            }
    }
}

// Units per tile (none for a tile without bin entries: tsc_rows_k writes
// its zeros) and whether the tile needs the multi-unit side buffer.
__global__ void tsc_units_k(const unsigned *tcount, long long ntiles, unsigned unit,
                            unsigned *nunits, unsigned *multi) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    unsigned c = tcount[t];
    unsigned u = (c + unit - 1) / unit;
    nunits[t] = u;
    multi[t] = u > 1 ? 1u : 0u;
}

__global__ void tsc_unit_map_k(const unsigned *ustart, const unsigned *nunits,
                               long long ntiles, unsigned *unit_tile) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    unsigned s = ustart[t];
    for (unsigned j = 0; j < nunits[t]; j++) unit_tile[s + j] = (unsigned)t;
}

// Fixed-point scale of each halo from its max|m| and selected count.
__global__ void tsc_scale_k(const unsigned *nsel, const unsigned long long *maxm_bits, int nhalo,
                            double *scale) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nhalo) return;
    double maxm = __longlong_as_double((long long)maxm_bits[h]);
    double bound = maxm * (double)nsel[h] * (1.0 + 1.0 / 1024.0);
    int S = 0;
    if (bound > 0.0) {
        int e;
        frexp(bound, &e);  // bound < 2^e
        S = 62 - e;
    }
    double vf = 0.001 / 0.004;  // (0.001/TSC_CELL_SIZE)^3, tsc_omp2.c:188
    vf = vf * vf * vf;
    scale[2 * h] = ldexp(1.0, S);
    scale[2 * h + 1] = ldexp(1.0, -S) * vf;
}

// The rows of tile t (TZ runs of TY * mx consecutive floats, clipped to the
// grid): value2(l) gives the tile's nodes l and l + 1 (x + mx (ly + TY lz)
// order).  Written as float pairs: mx and every grid offset are even and the
// grid is 8-byte aligned, so each pair is an aligned 8-byte store.
template <class F2>
__device__ inline void write_tile_rows(float *den, const HaloSeg &sg, const TileGeom &tg,
                                       unsigned tl, F2 value2) {
    const int ty = (int)(tl % (unsigned)tg.nty), tz = (int)(tl / (unsigned)tg.nty);
    const long long mx = tg.mx;
    const int nrows = min(tg.TY, sg.g.ny - ty * tg.TY);
    const int run2 = nrows * tg.mx / 2;
    for (int lz = 0; lz < tg.TZ; lz++) {
        const long long z = (long long)tz * tg.TZ + lz;
        if (z >= sg.g.nz) break;
        float2 *dst = reinterpret_cast<float2 *>(
            den + sg.goff + mx * ((long long)ty * tg.TY + (long long)sg.g.ny * z));
        const int l0 = lz * tg.TY * tg.mx;
        for (int i = threadIdx.x; i < run2; i += blockDim.x) dst[i] = value2(l0 + 2 * i);
    }
}

// Nodes l, l + 1 of a fixed-point tile as floats; a pair of empty nodes
// (most of a sparse tile) skips the conversion.
__device__ inline float2 fixed_pair(const unsigned long long *acc, int l, double inv) {
    const unsigned long long a0 = acc[l], a1 = acc[l + 1];
    float2 v = make_float2(0.0f, 0.0f);
    if ((a0 | a1) != 0ull) {
        v.x = (float)((double)(long long)a0 * inv);
        v.y = (float)((double)(long long)a1 * inv);
    }
    return v;
}

template <class P>
__global__ void __launch_bounds__(kBlock)
tsc_deposit_k(const HaloSeg *segs, HaloSeg seg0, const unsigned *tile_halo,
              const unsigned *tstart, const unsigned *tcount, const unsigned *ustart,
              const unsigned *nunits, const unsigned *unit_tile, const unsigned *total_units,
              const unsigned *multi_slot, const BinRec<P> *bins, const double *scale,
              float *den, unsigned long long *side, unsigned unit) {
    __shared__ unsigned long long acc[kTileCap];
    const unsigned nunit_total = *total_units;
    for (unsigned u = blockIdx.x; u < nunit_total; u += gridDim.x) {
        unsigned t = unit_tile[u];
        const int h = tile_halo ? (int)tile_halo[t] : 0;
        const HaloSeg sg = seg_at(segs, seg0, h);
        const Grid g = sg.g;
        const TileGeom tg = tile_geom(g);
        const double sc = scale[2 * h], inv = scale[2 * h + 1];
        unsigned chunk = u - ustart[t];
        unsigned nu = nunits[t];
        const unsigned tl = (unsigned)(t - sg.toff);
        const long long oy = (long long)(tl % (unsigned)tg.nty) * tg.TY;
        const long long oz = (long long)(tl / (unsigned)tg.nty) * tg.TZ;
        const int nodes = tg.nodes();
        for (int i = threadIdx.x; i < nodes; i += kBlock) acc[i] = 0ull;
        __syncthreads();
        unsigned e0 = tstart[t] + chunk * unit;
        unsigned e1 = min(e0 + unit, tstart[t] + tcount[t]);
        for (unsigned e = e0 + threadIdx.x; e < e1; e += kBlock) {
            const BinRec<P> rec = bins[e];
            Axis ax[3];
            rec_axes(rec, g, ax);
#pragma unroll
            for (int c = -1; c <= 1; c++) {
                long long nzn = ax[2].nr + c - oz;
                if (nzn < 0 || nzn >= tg.TZ) continue;
                float wz = axis_weight(ax[2], c);
#pragma unroll
                for (int b = -1; b <= 1; b++) {
                    long long nyn = ax[1].nr + b - oy;
                    if (nyn < 0 || nyn >= tg.TY) continue;
                    float wy = axis_weight(ax[1], b);
                    const long long row = (long long)tg.mx * (nyn + (long long)tg.TY * nzn);
#pragma unroll
                    for (int a = -1; a <= 1; a++) {
                        // x nodes stay in [0, mx) (particle_axes checked at count)
                        const long long nxn = ax[0].nr + a;
                        float wx = axis_weight(ax[0], a);
                        float wxy = wx * wy;
                        float cval = wxy * wz;
                        long long q = (long long)rint((double)cval * sc);
                        atomicAdd(&acc[nxn + row], (unsigned long long)q);
                    }
                }
            }
        }
        __syncthreads();
        if (nu == 1) {
            // the tile's rows, straight into the grid
            write_tile_rows(den, sg, tg, tl, [&](int l) { return fixed_pair(acc, l, inv); });
        } else {
            unsigned long long *dst = side + (size_t)multi_slot[t] * kTileCap;
            for (int l = threadIdx.x; l < nodes; l += kBlock)
                if (acc[l]) atomicAdd(&dst[l], acc[l]);
        }
        __syncthreads();
    }
}

// The rows of the tiles the deposit did not write: zeros for tiles without
// bin entries, the side-buffer sums for multi-unit tiles.  Grid-stride over
// the tiles, one workgroup per tile.
__global__ void __launch_bounds__(256)
tsc_rows_k(const HaloSeg *segs, HaloSeg seg0, const unsigned *tile_halo, long long ntiles,
           const unsigned *nunits, const unsigned *multi_slot,
           const unsigned long long *side, const double *scale, float *den) {
    for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const unsigned nu = nunits[t];
        if (nu == 1) continue;
        const int h = tile_halo ? (int)tile_halo[t] : 0;
        const HaloSeg sg = seg_at(segs, seg0, h);
        const TileGeom tg = tile_geom(sg.g);
        const unsigned tl = (unsigned)(t - sg.toff);
        if (nu == 0) {
            write_tile_rows(den, sg, tg, tl, [](int) { return make_float2(0.0f, 0.0f); });
        } else {
            const double inv = scale[2 * h + 1];
            const unsigned long long *src = side + (size_t)multi_slot[t] * kTileCap;
            write_tile_rows(den, sg, tg, tl, [&](int l) { return fixed_pair(src, l, inv); });
        }
    }
}

// Zeroes up to five word spans in one launch (the deposit's counters).
struct ZeroSpans {
    unsigned *p[5];
    long long n[5];
};
__global__ void zero_spans_k(ZeroSpans z) {
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long st = (long long)gridDim.x * blockDim.x;
#pragma unroll
    for (int k = 0; k < 5; k++)
        for (long long i = i0; i < z.n[k]; i += st) z.p[k][i] = 0u;
}

// Zero the side slots actually used (their count is known on device only).
__global__ void fill_side_k(unsigned long long *p, const unsigned *nmulti) {
    long long n = (long long)(*nmulti) * kTileCap;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// ---- read-back -------------------------------------------------------------
struct RAxis {
    int nr;
    long sg;
    float w1, w2, w3;
};

__device__ inline void readback_axis(double p, RAxis &a) {
    a.nr = (int)rint(p);
    float off = (float)(p - (double)a.nr);
    a.sg = (long)copysignf(1.0f, off);
    float d = (float)fabs((double)off);
    float d2 = d * d;
    a.w1 = (float)(0.75 - (double)d2);
    a.w3 = (float)(0.5 * (0.25 + (double)d * ((double)d - 1.0)));
    a.w2 = a.w3 + d;
}

template <class P>
__global__ void __launch_bounds__(kBlock)
tsc_readback_k(const P *x, const P *y, const P *z, const HaloSeg *segs, HaloSeg seg0,
               const PChunk *chunks, int per, const float *den, float *out, unsigned *status,
               const unsigned *order, const long long *shift) {
    const PChunk ch = chunk_at(chunks, seg0, per);
    const HaloSeg sg = seg_at(segs, seg0, ch.halo);
    const Grid g = sg.g;
    const float *hden = den + sg.goff;
    const long mx = (long)g.mx();
    // particle i's coordinates sit at p = i + sh, its result at order[p]
    const long long sh = shift ? shift[ch.halo] : 0;
    for (long long i = ch.b0 + threadIdx.x; i < ch.b1; i += kBlock) {
        const long long p = i + sh;
        const long long o = order ? (long long)order[p] : i;
        RAxis a[3];
        readback_axis(((double)x[p] - g.x0) / g.h, a[0]);
        readback_axis(((double)y[p] - g.y0) / g.h, a[1]);
        readback_axis(((double)z[p] - g.z0) / g.h, a[2]);
        if (a[0].nr < 1 || a[0].nr > g.mx() - 2 || a[1].nr < 1 || a[1].nr > g.ny - 2 ||
            a[2].nr < 1 || a[2].nr > g.nz - 2) {
            atomicOr(status, (unsigned)DEV_ERR_OUT_OF_GRID);
            out[o] = 0.0f;
            continue;
        }
        long ix[3] = {a[0].nr, a[0].nr + a[0].sg, a[0].nr - a[0].sg};
        long iy[3] = {a[1].nr, a[1].nr + a[1].sg, a[1].nr - a[1].sg};
        long iz[3] = {a[2].nr, a[2].nr + a[2].sg, a[2].nr - a[2].sg};
        float wx[3] = {a[0].w1, a[0].w2, a[0].w3};
        float wy[3] = {a[1].w1, a[1].w2, a[1].w3};
        float wz[3] = {a[2].w1, a[2].w2, a[2].w3};
        float wxy[3][3];
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < 3; c++) wxy[b][c] = wx[c] * wy[b];
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int b = 0; b < 3; b++)
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    float gv = hden[ix[q] + mx * (iy[b] + (long)g.ny * iz[c])];
                    s += (gv * wxy[b][q]) * wz[c];
                }
        out[o] = s;
    }
}

// ---- adjoint (no counterpart in the reference) -----------------------------
// One axis in FP64: nodes (nearest rint(p), the offset's side, the other
// side), weights w(u) = 0.75 - u^2 (|u| <= 0.5), 0.5 (1.5 - |u|)^2 (0.5 <
// |u| < 1.5) of u = p - node, and their derivatives dw/dp.
__device__ inline void grad_axis(double p, long id[3], double w[3], double dw[3]) {
    const long nr = (long)rint(p);
    const double off = p - (double)nr;
    const double sg = off < 0.0 ? -1.0 : 1.0;
    const double d = fabs(off);
    id[0] = nr;
    id[1] = nr + (long)sg;
    id[2] = nr - (long)sg;
    w[0] = 0.75 - d * d;
    dw[0] = -2.0 * off;
    w[1] = 0.5 * (0.5 + d) * (0.5 + d);
    dw[1] = (0.5 + d) * sg;
    w[2] = 0.5 * (0.5 - d) * (0.5 - d);
    dw[2] = -(0.5 - d) * sg;
}

// Gradient read-back of L(rho_tsc), rho_tsc = R S D: with gin = dL/drho,
// F the forward smoothed grid and A = S D_gin the adjoint grid,
//   dL/dm_p = sum_nodes W A,   dL/dx_p = (1/h) sum_nodes dW/dx (g_p F + m_p A).
// acc holds [gx | gy | gz | gm] (nacc each, FP64); this kernel adds its part.
// One block per particle chunk (batch) or per `per` particles (one halo).
template <class P>
__global__ void __launch_bounds__(kBlock)
tsc_gradback_k(const P *x, const P *y, const P *z, const P *m, const float *gin,
               const HaloSeg *segs, HaloSeg seg0, const PChunk *chunks, int per,
               const float *F, const float *A, double *acc, long long nacc,
               unsigned *status) {
    const PChunk ch = chunk_at(chunks, seg0, per);
    const HaloSeg sg = seg_at(segs, seg0, ch.halo);
    const Grid g = sg.g;
    const float *hF = F + sg.goff, *hA = A + sg.goff;
    const long mx = (long)g.mx();
    const long n[3] = {mx, g.ny, g.nz};
    for (long long p = ch.b0 + threadIdx.x; p < ch.b1; p += kBlock) {
        long id[3][3];
        double w[3][3], dw[3][3];
        grad_axis(((double)x[p] - g.x0) / g.h, id[0], w[0], dw[0]);
        grad_axis(((double)y[p] - g.y0) / g.h, id[1], w[1], dw[1]);
        grad_axis(((double)z[p] - g.z0) / g.h, id[2], w[2], dw[2]);
        bool inside = true;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int q = 0; q < 3; q++) inside &= id[a][q] >= 0 && id[a][q] < n[a];
        if (!inside) {
            atomicOr(status, (unsigned)DEV_ERR_OUT_OF_GRID);
            continue;
        }
        const double gp = (double)gin[p], mp = (double)m[p];
        double sx = 0.0, sy = 0.0, sz = 0.0, sm = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int b = 0; b < 3; b++)
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const long o = id[0][a] + mx * (id[1][b] + (long)g.ny * id[2][c]);
                    const double fa = (double)hA[o];
                    const double v = gp * (double)hF[o] + mp * fa;
                    sx += dw[0][a] * w[1][b] * w[2][c] * v;
                    sy += w[0][a] * dw[1][b] * w[2][c] * v;
                    sz += w[0][a] * w[1][b] * dw[2][c] * v;
                    sm += w[0][a] * w[1][b] * w[2][c] * fa;
                }
        acc[p] += sx / g.h;
        acc[nacc + p] += sy / g.h;
        acc[2 * nacc + p] += sz / g.h;
        acc[3 * nacc + p] += sm;
    }
}

// ---- peak climbing (no counterpart in the reference) -------------------------
// Gradient (cell units) of the read-back field of the smoothed grid F at cell
// coordinate q, FP64 TSC weights; false when the stencil leaves the grid.
__device__ inline bool field_grad(const double q[3], const Grid &g, const float *F,
                                  double G[3]) {
    long id[3][3];
    double w[3][3], dw[3][3];
    const long n[3] = {(long)g.nx, g.ny, g.nz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        grad_axis(q[a], id[a], w[a], dw[a]);
#pragma unroll
        for (int t = 0; t < 3; t++)
            if (id[a][t] < 0 || id[a][t] >= n[a]) return false;
    }
    const long mx = (long)g.mx();
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const double f = (double)F[id[0][a] + mx * (id[1][b] + (long)g.ny * id[2][c])];
                gx += dw[0][a] * w[1][b] * w[2][c] * f;
                gy += w[0][a] * dw[1][b] * w[2][c] * f;
                gz += w[0][a] * w[1][b] * dw[2][c] * f;
            }
    G[0] = gx;
    G[1] = gy;
    G[2] = gz;
    return true;
}

// Peak climbing: niter normalised gradient-ascent steps of `step` cells on
// the read-back field of the smoothed grid (positions clamped to [1, n-2]
// cells so every stencil stays inside), then membership: from the nearest
// node (rint) an exact integer steepest ascent over the 26-neighbourhood of
// F (move to the largest neighbour while it is strictly larger than the
// current node; among equal neighbours the lowest index) to a local maximum,
// whose unpadded flat index i + nx (j + ny k) is the particle's peak.  Each
// move strictly increases F, so the walk ends; kClimbMax bounds it.
constexpr int kClimbMax = 4096;

template <class P>
__global__ void __launch_bounds__(kBlock)
peak_climb_k(const P *x, const P *y, const P *z, const HaloSeg *segs, HaloSeg seg0,
             const PChunk *chunks, int per, const float *den, int niter, double step,
             float *ox, float *oy, float *oz, int *peak) {
    const PChunk ch = chunk_at(chunks, seg0, per);
    const HaloSeg sg = seg_at(segs, seg0, ch.halo);
    const Grid g = sg.g;
    const float *F = den + sg.goff;
    const long long mx = g.mx();
    const double lim[3] = {(double)(g.nx - 2), (double)(g.ny - 2), (double)(g.nz - 2)};
    for (long long p = ch.b0 + threadIdx.x; p < ch.b1; p += kBlock) {
        double q[3] = {((double)x[p] - g.x0) / g.h, ((double)y[p] - g.y0) / g.h,
                       ((double)z[p] - g.z0) / g.h};
#pragma unroll
        for (int a = 0; a < 3; a++) q[a] = fmin(fmax(q[a], 1.0), lim[a]);
        for (int it = 0; it < niter; it++) {
            double G[3];
            if (!field_grad(q, g, F, G)) break;
            const double nrm = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
            if (!(nrm > 0.0)) break;
#pragma unroll
            for (int a = 0; a < 3; a++) q[a] = fmin(fmax(q[a] + step * G[a] / nrm, 1.0), lim[a]);
        }
        ox[p] = (float)(g.x0 + q[0] * g.h);
        oy[p] = (float)(g.y0 + q[1] * g.h);
        oz[p] = (float)(g.z0 + q[2] * g.h);
        long c[3];
#pragma unroll
        for (int a = 0; a < 3; a++) c[a] = (long)rint(q[a]);
        float fc = F[c[0] + mx * (c[1] + (long long)g.ny * c[2])];
        const int nx = g.nx, ny = g.ny, nz = g.nz, mxi = (int)mx;
        for (int step_i = 0; step_i < kClimbMax; step_i++) {
            long best[3] = {c[0], c[1], c[2]};
            float fb = fc;
            if (c[0] >= 1 && c[0] <= nx - 2 && c[1] >= 1 && c[1] <= ny - 2 && c[2] >= 1 &&
                c[2] <= nz - 2) {
                // all 26 neighbours inside: their values first (independent
                // loads, 32-bit offsets from the cell), then the same scan in
                // flat-index order with the same strict comparison
                const float *Fc = F + (c[0] + mx * (c[1] + (long long)ny * c[2]));
                float f[27];
#pragma unroll
                for (int dz = -1; dz <= 1; dz++)
#pragma unroll
                    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
                        for (int dx = -1; dx <= 1; dx++)
                            f[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1] =
                                Fc[dx + mxi * (dy + ny * dz)];
                int bi = 13;  // the cell itself
#pragma unroll
                for (int t = 0; t < 27; t++)
                    if (f[t] > fb) {
                        fb = f[t];
                        bi = t;
                    }
                best[0] = c[0] + bi % 3 - 1;
                best[1] = c[1] + (bi / 3) % 3 - 1;
                best[2] = c[2] + bi / 9 - 1;
            } else {
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const long i = c[0] + dx, j = c[1] + dy, k = c[2] + dz;
                            if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz)
                                continue;
                            const float f = F[i + mx * (j + (long long)ny * k)];
                            // loop order visits neighbours by increasing flat
                            // index: a strictly larger value wins, so ties keep
                            // the lowest
                            if (f > fb) {
                                fb = f;
                                best[0] = i;
                                best[1] = j;
                                best[2] = k;
                            }
                        }
            }
            if (!(fb > fc)) break;
            c[0] = best[0];
            c[1] = best[1];
            c[2] = best[2];
            fc = fb;
        }
        peak[p] = (int)(c[0] + (long long)g.nx * (c[1] + (long long)g.ny * c[2]));
    }
}

HaloSeg single_seg(const Grid &g, long long np) {
    HaloSeg sg;
    sg.p0 = 0;
    sg.np = np;
    sg.goff = 0;
    sg.toff = 0;
    sg.g = g;
    return sg;
}

int cu_count(int device) {
    static int ncu = 0;
    if (ncu == 0) {
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) !=
                hipSuccess || ncu <= 0)
            ncu = 256;
    }
    return ncu;
}

// Deposit of the halos of `segs` (or the single halo `seg0`) through the
// tile bins.  ntiles, the chunk count and max_halo_tiles describe the tile
// space and the particle blocks; tile_halo maps a tile to its halo.
template <class P>
hipError_t deposit_impl(const P *x, const P *y, const P *z, const P *m, const int *type,
                        int ptype, int nhalo, const HaloSeg *segs, const HaloSeg &seg0,
                        const PChunk *chunks, long long nchunks, const unsigned *tile_halo,
                        long long ntiles, long long max_halo_tiles, int max_entries, long long np,
                        float *den, const Schedule &sc, Workspace &ws, hipStream_t s) {
    const unsigned unit = (unsigned)sc.tsc_unit_entries;
    WsScope scope(ws);  // bins and side buffer end with the deposit
    int use_lds = max_halo_tiles <= kLdsHistMax ? 1 : 0;
    unsigned *tcount = ws.alloc_n<unsigned>(ntiles);
    unsigned *tcursor = ws.alloc_n<unsigned>(ntiles);
    unsigned *tstart = ws.alloc_n<unsigned>(ntiles);
    unsigned *nunits = ws.alloc_n<unsigned>(ntiles);
    unsigned *ustart = ws.alloc_n<unsigned>(ntiles);
    unsigned *multi = ws.alloc_n<unsigned>(ntiles);
    unsigned *misc = ws.alloc_n<unsigned>(8);  // [1]=total entries [2]=units [3]=nmulti
    unsigned *nsel = ws.alloc_n<unsigned>(nhalo);
    unsigned long long *maxm = ws.alloc_n<unsigned long long>(nhalo);
    double *scale = ws.alloc_n<double>(2 * (size_t)nhalo);
    if (!tcount || !tcursor || !tstart || !nunits || !ustart || !multi || !misc || !nsel ||
        !maxm || !scale)
        return hipErrorOutOfMemory;
    {
        ZeroSpans z{{tcount, tcursor, misc, nsel, reinterpret_cast<unsigned *>(maxm)},
                    {ntiles, ntiles, 8, nhalo, 2LL * nhalo}};
        const long long mx = std::max<long long>(ntiles, 2LL * nhalo);
        const unsigned zb = (unsigned)std::min<long long>((mx + 255) / 256, 1024);
        zero_spans_k<<<zb, 256, 0, s>>>(z);
    }
    size_t lds_count = use_lds ? sizeof(unsigned) * max_halo_tiles : 0;
    if (nchunks > 0)
        tsc_count_k<P><<<dim3((unsigned)nchunks), dim3(kBlock), lds_count, s>>>(
            x, y, z, m, type, ptype, segs, seg0, chunks, tcount, nsel, maxm, ws.status(),
            use_lds);
    PGALF_HIP_TRY(exclusive_scan_u32(tcount, tstart, ntiles, &misc[1], ws, s));
    // at most max_entries tiles per particle (2 x 2 unless a tile is one row)
    BinRec<P> *bins = ws.alloc_n<BinRec<P>>((size_t)np * max_entries + 1);
    if (!bins) return hipErrorOutOfMemory;
    if (nchunks > 0)
        tsc_scatter_k<P><<<dim3((unsigned)nchunks), dim3(kBlock), 2 * lds_count, s>>>(
            x, y, z, m, type, ptype, segs, seg0, chunks, tstart, tcursor, bins, use_lds);
    unsigned tb = (unsigned)((ntiles + 255) / 256);
    tsc_units_k<<<tb, 256, 0, s>>>(tcount, ntiles, unit, nunits, multi);
    PGALF_HIP_TRY(exclusive_scan_u32(nunits, ustart, ntiles, &misc[2], ws, s));
    PGALF_HIP_TRY(exclusive_scan_u32(multi, multi, ntiles, &misc[3], ws, s));
    // unit -> tile map: the unit count is bounded by ntiles + entries/unit
    const size_t max_ent = (size_t)np * max_entries;
    size_t max_units = (size_t)ntiles + max_ent / unit + 1;
    unsigned *unit_tile = ws.alloc_n<unsigned>(max_units);
    // multi-unit side buffer: a tile with > unit entries needs one slot;
    // there are at most entries/unit of them.
    size_t max_multi = max_ent / unit + 1;
    if ((long long)max_multi > ntiles) max_multi = (size_t)ntiles;
    unsigned long long *side = ws.alloc_n<unsigned long long>(max_multi * kTileCap);
    if (!unit_tile || !side) return hipErrorOutOfMemory;
    tsc_unit_map_k<<<tb, 256, 0, s>>>(ustart, nunits, ntiles, unit_tile);
    tsc_scale_k<<<(unsigned)((nhalo + 255) / 256), 256, 0, s>>>(nsel, maxm, nhalo, scale);
    const int ncu = cu_count(ws.device());
    fill_side_k<<<(unsigned)ncu * 4, 256, 0, s>>>(side, &misc[3]);
    tsc_deposit_k<P><<<(unsigned)ncu * 4, kBlock, 0, s>>>(
        segs, seg0, tile_halo, tstart, tcount, ustart, nunits, unit_tile, &misc[2], multi, bins,
        scale, den, side, unit);
    const long long rb = ntiles < (long long)ncu * 8 ? ntiles : (long long)ncu * 8;
    if (rb > 0)
        tsc_rows_k<<<(unsigned)rb, 256, 0, s>>>(segs, seg0, tile_halo, ntiles, nunits, multi, side,
                                                scale, den);
    return hipGetLastError();
}

}  // namespace

long long tsc_tile_count(const Grid &g) { return tile_geom(g).ntiles(); }

bool tsc_grid_fits(const Grid &g) { return g.mx() <= kTileCap; }

int tsc_tile_entries(const Grid &g) {
    const TileGeom tg = tile_geom(g);
    return (tg.TY == 1 ? 3 : 2) * (tg.TZ == 1 ? 3 : 2);
}

template <class P>
hipError_t tsc_deposit(const P *x, const P *y, const P *z, const P *m, const int *type,
                       long long np, int ptype, const Grid &g, float *den, Workspace &ws,
                       hipStream_t s) {
    const HaloSeg seg0 = single_seg(g, np);
    if (!tsc_grid_fits(g)) return hipErrorInvalidValue;
    const long long ntiles = tile_geom(g).ntiles();
    const long long nchunks = (np + kPerBlock - 1) / kPerBlock;
    return deposit_impl<P>(x, y, z, m, type, ptype, 1, nullptr, seg0, nullptr, nchunks, nullptr,
                           ntiles, ntiles, tsc_tile_entries(g), np, den, schedule_for(np), ws, s);
}

template <class P>
hipError_t tsc_deposit_batch(const P *x, const P *y, const P *z, const P *m, const BatchDev &b,
                             float *den, Workspace &ws, hipStream_t s) {
    static_assert(kPerBlock == kChunkParticles && kFineParticles == kBlock,
                  "batch chunks are deposit blocks; fine chunks are read-back blocks");
    const HaloSeg seg0{};
    return deposit_impl<P>(x, y, z, m, nullptr, PGALF_TYPE_ALL, b.nhalo, b.segs, seg0, b.chunks,
                           b.nchunks, b.tile_halo, b.ntiles, b.max_halo_tiles,
                           b.max_tile_entries, b.np, den, b.sched, ws, s);
}

template <class P>
hipError_t tsc_readback(const P *x, const P *y, const P *z, long long np, const Grid &g,
                        const float *den, float *out, Workspace &ws, hipStream_t s,
                        const unsigned *order) {
    if (np <= 0) return hipSuccess;
    const HaloSeg seg0 = single_seg(g, np);
    tsc_readback_k<P><<<(unsigned)((np + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        x, y, z, nullptr, seg0, nullptr, kBlock, den, out, ws.status(), order, nullptr);
    return hipGetLastError();
}

template <class P>
hipError_t tsc_readback_batch(const P *x, const P *y, const P *z, const BatchDev &b,
                              const float *den, float *out, Workspace &ws, hipStream_t s,
                              const unsigned *order, const long long *shift) {
    if (b.nfine <= 0) return hipSuccess;
    const HaloSeg seg0{};
    tsc_readback_k<P><<<(unsigned)b.nfine, kBlock, 0, s>>>(x, y, z, b.segs, seg0, b.fine,
                                                          kFineParticles, den, out, ws.status(),
                                                          order, shift);
    return hipGetLastError();
}

template <class P>
hipError_t tsc_grad_readback(const P *x, const P *y, const P *z, const P *m, const float *gin,
                             long long np, const Grid &g, const float *F, const float *A,
                             double *acc, Workspace &ws, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    const HaloSeg seg0 = single_seg(g, np);
    tsc_gradback_k<P><<<(unsigned)((np + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        x, y, z, m, gin, nullptr, seg0, nullptr, kBlock, F, A, acc, np, ws.status());
    return hipGetLastError();
}

template <class P>
hipError_t tsc_grad_readback_batch(const P *x, const P *y, const P *z, const P *m,
                                   const float *gin, const BatchDev &b, const float *F,
                                   const float *A, double *acc, Workspace &ws, hipStream_t s) {
    if (b.nfine <= 0) return hipSuccess;
    const HaloSeg seg0{};
    tsc_gradback_k<P><<<(unsigned)b.nfine, kBlock, 0, s>>>(
        x, y, z, m, gin, b.segs, seg0, b.fine, kFineParticles, F, A, acc, b.np, ws.status());
    return hipGetLastError();
}

template <class P>
hipError_t peak_climb(const P *x, const P *y, const P *z, long long np, const Grid &g,
                      const float *den, int niter, double step, float *ox, float *oy,
                      float *oz, int *peak, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    if (g.nx < 3 || g.ny < 3 || g.nz < 3) return hipErrorInvalidValue;
    const HaloSeg seg0 = single_seg(g, np);
    peak_climb_k<P><<<(unsigned)((np + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        x, y, z, nullptr, seg0, nullptr, kBlock, den, niter, step, ox, oy, oz, peak);
    return hipGetLastError();
}

template <class P>
hipError_t peak_climb_batch(const P *x, const P *y, const P *z, const BatchDev &b,
                            const float *den, int niter, double step, float *ox, float *oy,
                            float *oz, int *peak, hipStream_t s) {
    if (b.nfine <= 0) return hipSuccess;
    const HaloSeg seg0{};
    peak_climb_k<P><<<(unsigned)b.nfine, kBlock, 0, s>>>(x, y, z, b.segs, seg0, b.fine,
                                                        kFineParticles, den, niter, step, ox,
                                                        oy, oz, peak);
    return hipGetLastError();
}

#define PGALF_TSC_INST(P)                                                                      \
    template hipError_t tsc_grad_readback<P>(const P *, const P *, const P *, const P *,       \
                                             const float *, long long, const Grid &,           \
                                             const float *, const float *, double *,           \
                                             Workspace &, hipStream_t);                        \
    template hipError_t tsc_grad_readback_batch<P>(const P *, const P *, const P *, const P *, \
                                                   const float *, const BatchDev &,            \
                                                   const float *, const float *, double *,     \
                                                   Workspace &, hipStream_t);                  \
    template hipError_t peak_climb<P>(const P *, const P *, const P *, long long,              \
                                      const Grid &, const float *, int, double, float *,       \
                                      float *, float *, int *, hipStream_t);                   \
    template hipError_t peak_climb_batch<P>(const P *, const P *, const P *, const BatchDev &, \
                                            const float *, int, double, float *, float *,      \
                                            float *, int *, hipStream_t);                      \
    template hipError_t tsc_deposit<P>(const P *, const P *, const P *, const P *, const int *, \
                                       long long, int, const Grid &, float *, Workspace &,      \
                                       hipStream_t);                                           \
    template hipError_t tsc_readback<P>(const P *, const P *, const P *, long long,            \
                                        const Grid &, const float *, float *, Workspace &,     \
                                        hipStream_t, const unsigned *);                        \
    template hipError_t tsc_deposit_batch<P>(const P *, const P *, const P *, const P *,       \
                                             const BatchDev &, float *, Workspace &,           \
                                             hipStream_t);                                     \
    template hipError_t tsc_readback_batch<P>(const P *, const P *, const P *,                 \
                                              const BatchDev &, const float *, float *,        \
                                              Workspace &, hipStream_t, const unsigned *,      \
                                              const long long *);
PGALF_TSC_INST(float)
PGALF_TSC_INST(double)
#undef PGALF_TSC_INST

}  // namespace pgalf
