// Exact k-nearest-neighbour lists and SPH (W4) kNN densities for gfx950.
//
// Semantics (NewGalFinder/nnost.c): for every query particle, the k nearest
// candidates, self included, ordered by the FP64 squared distance formed as
// the reference forms it (query minus candidate, dx*dx+dy*dy+dz*dz, no FMA;
// nnost.c:362-365) -- and, where distances tie, by candidate index.  Then
// optionally rho_i = sum_j W4(r_ij, h_i) [* m_j] with h_i half the distance
// to the k-th neighbour (nnost.c:536-555, 607-623, 689-704).
//
// Structure:
//   * queries and candidates are sorted by 30-bit Morton code (stable radix
//     sort, ties by index) so a 64-particle bucket is spatially compact;
//   * candidate buckets get FP64 bounding boxes and an implicit binary tree
//     of boxes over consecutive buckets;
//   * one wavefront serves one query bucket, one query per lane.  The wave
//     walks the box tree near-first with a wave-uniform stack, pruning nodes
//     farther (box to box) than the largest current k-th distance of its
//     lanes; each surviving leaf bucket is staged in LDS and every lane scans
//     its 64 candidates.  Accepted candidates go to a per-lane LDS buffer;
//     when any lane's buffer may overflow, all lanes merge it into their
//     sorted top-32 list held in registers (bitonic sort of 16 + bitonic
//     merge), tightening their k-th distance.
//   * keys are (FP64 dist2, uint32 index) compared lexicographically, so the
//     lists are exact and deterministic regardless of visiting order.
// Box pruning is conservative: box gaps are formed with the same monotone
// FP64 operations as the distances, from box faces that are member
// coordinates, so a gap never exceeds a member's computed distance.
#include <cfloat>

#include "pgalf_internal.h"

namespace pgalf {
namespace {

constexpr int kWave = 64;
constexpr int kTop = 32;     // list capacity (k <= 32, MAX_NUM_NEAR tree.h:20)
constexpr int kBuf = 16;     // per-lane LDS buffer
constexpr int kChunkC = 8;   // candidates between overflow checks
constexpr int kMaxLevels = 32;
constexpr int kStack = 96;

struct Box {
    double lo[3], hi[3];
};

// Implicit box tree over nb leaf buckets: level l holds ceil(nb / 2^l)
// nodes, stored after all lower levels.
struct Levels {
    int nlev;
    long long nb;
};

__host__ __device__ inline long long level_count(long long nb, int l) {
    return (nb + (1LL << l) - 1) >> l;
}
__host__ __device__ inline long long level_offset(long long nb, int l) {
    long long o = 0;
    for (int i = 0; i < l; i++) o += level_count(nb, i);
    return o;
}

struct Key {
    double d;
    unsigned i;
};

// Branch-free comparisons: short-circuit forms turn into control flow that
// the compiler then folds into pointer selects, pinning the key arrays to
// scratch memory.
__device__ inline bool klt(double ad, unsigned ai, double bd, unsigned bi) {
    return (ad < bd) | ((ad == bd) & (ai < bi));
}
__device__ inline bool key_lt(const Key &a, const Key &b) { return klt(a.d, a.i, b.d, b.i); }

__device__ inline void cas2(double &ad, unsigned &ai, double &bd, unsigned &bi) {
    bool sw = klt(bd, bi, ad, ai);  // a <= b afterwards
    double t0 = ad, t2 = bd;
    unsigned t1 = ai, t3 = bi;
    ad = sw ? t2 : t0;
    ai = sw ? t3 : t1;
    bd = sw ? t0 : t2;
    bi = sw ? t1 : t3;
}

__device__ inline double d2_ref(double qx, double qy, double qz, double cx, double cy,
                                double cz) {
    double tx = qx - cx, ty = qy - cy, tz = qz - cz;
    double a = tx * tx;
    double b = ty * ty;
    double c = tz * tz;
    return (a + b) + c;
}

__device__ inline double gap1(double lo, double hi, double qlo, double qhi) {
    // smallest |c - q| for c in [lo,hi], q in [qlo,qhi]
    if (lo > qhi) return lo - qhi;
    if (qlo > hi) return qlo - hi;
    return 0.0;
}

__device__ inline double box_gap2(const Box &b, const double qlo[3], const double qhi[3]) {
    double g0 = gap1(b.lo[0], b.hi[0], qlo[0], qhi[0]);
    double g1 = gap1(b.lo[1], b.hi[1], qlo[1], qhi[1]);
    double g2 = gap1(b.lo[2], b.hi[2], qlo[2], qhi[2]);
    double a = g0 * g0;
    double bb = g1 * g1;
    double c = g2 * g2;
    return (a + bb) + c;
}

__device__ inline double wave_max_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_min_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// W4 cubic spline, arithmetic of nnost.c:536-555.
__device__ inline float W4(float dist, float h) {
    float r = dist / h;
    if ((double)r < 1.0) {
        double t = (2.0 - (double)r) * (double)r * (double)r - 1.333333;
        return (float)((-0.2387324) * t / (double)(h * h * h));
    } else if ((double)r < 2.0) {
        float twomr = (float)((2.0 - (double)r) / (double)h);
        return (float)((0.07957747) * (double)twomr * (double)twomr * (double)twomr);
    }
    return 0.0f;
}

__device__ inline unsigned long long ord_d(double v) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double unord_d(unsigned long long o) {
    unsigned long long b = (o >> 63) ? (o & 0x7fffffffffffffffull) : ~o;
    return __longlong_as_double((long long)b);
}

template <class P>
__global__ void bbox_k(const P *x, const P *y, const P *z, long long np,
                       unsigned long long *mm /* 3 min, 3 max (ordered) */) {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < np;
         i += (long long)gridDim.x * blockDim.x) {
        double v[3] = {(double)x[i], (double)y[i], (double)z[i]};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = fmin(lo[a], v[a]);
            hi[a] = fmax(hi[a], v[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = wave_min_d(lo[a]);
        hi[a] = wave_max_d(hi[a]);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            atomicMin(&mm[a], ord_d(lo[a]));
            atomicMax(&mm[3 + a], ord_d(hi[a]));
        }
    }
}

__device__ inline unsigned spread3(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Morton keys of particles idx[0..n) (idx == nullptr: identity).
template <class P>
__global__ void morton_k(const P *x, const P *y, const P *z, const unsigned *idx, long long n,
                         const unsigned long long *mm, unsigned *keys, unsigned *vals) {
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double lo[3], ext = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = unord_d(mm[a]);
        ext = fmax(ext, unord_d(mm[3 + a]) - lo[a]);
    }
    double sc = ext > 0.0 ? 1024.0 / ext : 0.0;
    unsigned i = idx ? idx[s] : (unsigned)s;
    double v[3] = {(double)x[i], (double)y[i], (double)z[i]};
    unsigned q[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double t = (v[a] - lo[a]) * sc;
        int qi = (int)t;
        q[a] = (unsigned)(qi < 0 ? 0 : (qi > 1023 ? 1023 : qi));
    }
    keys[s] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    vals[s] = i;
}

template <class P>
__global__ void gather_sorted_k(const P *x, const P *y, const P *z, const unsigned *perm,
                                long long n, double *sx, double *sy, double *sz) {
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    unsigned i = perm[s];
    sx[s] = (double)x[i];
    sy[s] = (double)y[i];
    sz[s] = (double)z[i];
}

__global__ void leaf_boxes_k(const double *sx, const double *sy, const double *sz, long long n,
                             Box *boxes) {
    long long b = blockIdx.x;
    long long i = b * kWave + threadIdx.x;
    bool v = i < n;
    double c[3] = {v ? sx[i] : DBL_MAX, v ? sy[i] : DBL_MAX, v ? sz[i] : DBL_MAX};
    double d[3] = {v ? sx[i] : -DBL_MAX, v ? sy[i] : -DBL_MAX, v ? sz[i] : -DBL_MAX};
    Box bx;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        bx.lo[a] = wave_min_d(c[a]);
        bx.hi[a] = wave_max_d(d[a]);
    }
    if (threadIdx.x == 0) boxes[b] = bx;
}

__global__ void parent_boxes_k(const Box *child, long long nchild, Box *parent,
                               long long nparent) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nparent) return;
    Box a = child[2 * j];
    if (2 * j + 1 < nchild) {
        Box b = child[2 * j + 1];
#pragma unroll
        for (int t = 0; t < 3; t++) {
            a.lo[t] = fmin(a.lo[t], b.lo[t]);
            a.hi[t] = fmax(a.hi[t], b.hi[t]);
        }
    }
    parent[j] = a;
}

// selected candidates -> flag
__global__ void select_flags_k(const int *type, long long np, int ptype, unsigned *flags) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) flags[i] = type[i] == ptype ? 1u : 0u;
}
__global__ void select_compact_k(const unsigned *flags, const unsigned *pos, long long np,
                                 unsigned *out) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np && flags[i]) out[pos[i]] = (unsigned)i;
}

// Bitonic networks written as compile-time recursions over split (dist,
// index) arrays so every index is a constant and the lists stay in
// registers.
template <int N, int K, int J>
__device__ inline void bitonic_sort_stage(double (&d)[N], unsigned (&x)[N]) {
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int l = i ^ J;
        if (l > i) {
            if ((i & K) == 0) cas2(d[i], x[i], d[l], x[l]);
            else cas2(d[l], x[l], d[i], x[i]);
        }
    }
    if constexpr (J > 1) bitonic_sort_stage<N, K, J / 2>(d, x);
}
template <int N, int K>
__device__ inline void bitonic_sort(double (&d)[N], unsigned (&x)[N]) {
    bitonic_sort_stage<N, K, K / 2>(d, x);
    if constexpr (K < N) bitonic_sort<N, 2 * K>(d, x);
}
template <int N, int J>
__device__ inline void bitonic_clean(double (&d)[N], unsigned (&x)[N]) {
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int l = i ^ J;
        if (l > i) cas2(d[i], x[i], d[l], x[l]);
    }
    if constexpr (J > 1) bitonic_clean<N, J / 2>(d, x);
}

// Merge a buffer of up to 16 keys into the sorted top-32: sort the buffer,
// take the elementwise minimum against its reverse (padded with +inf) --
// the 32 smallest, as a bitonic sequence -- and clean.
__device__ inline void merge_top(double (&td)[kTop], unsigned (&ti)[kTop],
                                 double (&bd)[kBuf], unsigned (&bi)[kBuf]) {
    bitonic_sort<kBuf, 2>(bd, bi);
#pragma unroll
    for (int j = 0; j < kBuf; j++) {
        double cd = bd[kBuf - 1 - j];
        unsigned ci = bi[kBuf - 1 - j];
        bool take = klt(cd, ci, td[kBuf + j], ti[kBuf + j]);
        td[kBuf + j] = take ? cd : td[kBuf + j];
        ti[kBuf + j] = take ? ci : ti[kBuf + j];
    }
    bitonic_clean<kTop, kTop / 2>(td, ti);
}

// k-th smallest of the sorted list as a masked maximum (a select on
// t == k-1 would be folded into a dynamic index).
__device__ inline Key kth_of(const double (&td)[kTop], const unsigned (&ti)[kTop], int k) {
    Key r;
    r.d = td[0];
    r.i = ti[0];
#pragma unroll
    for (int t = 1; t < kTop; t++) {
        bool take = (t < k) & klt(r.d, r.i, td[t], ti[t]);
        r.d = take ? td[t] : r.d;
        r.i = take ? ti[t] : r.i;
    }
    return r;
}

// Load a lane's buffer (cnt entries, rest +inf) and merge it.
__device__ inline void flush_buffer(double (&td)[kTop], unsigned (&ti)[kTop], int cnt,
                                    const double (*s_bd)[kWave], const unsigned (*s_bi)[kWave],
                                    int lane) {
    double bd[kBuf];
    unsigned bi[kBuf];
#pragma unroll
    for (int t = 0; t < kBuf; t++) {
        bool v = t < cnt;
        double dv = s_bd[t][lane];
        unsigned iv = s_bi[t][lane];
        bd[t] = v ? dv : __longlong_as_double(0x7ff0000000000000ll);
        bi[t] = v ? iv : 0xffffffffu;
    }
    merge_top(td, ti, bd, bi);
}

template <class P>
__global__ void __launch_bounds__(kWave)
knn_query_k(const double *qx, const double *qy, const double *qz, const unsigned *qid,
            long long nq, const double *cx, const double *cy, const double *cz,
            const unsigned *cid, long long nc, const Box *boxes, Levels lv, int k,
            long long *nbr, float *rho, int rho_mode, const P *mass) {
    __shared__ double s_x[kWave], s_y[kWave], s_z[kWave];
    __shared__ unsigned s_id[kWave];
    __shared__ double s_bd[kBuf][kWave];
    __shared__ unsigned s_bi[kBuf][kWave];
    __shared__ int s_stack[kStack];
    __shared__ long long s_off[kMaxLevels], s_n[kMaxLevels];

    const int lane = threadIdx.x;
    if (lane < lv.nlev) {
        s_off[lane] = level_offset(lv.nb, lane);
        s_n[lane] = level_count(lv.nb, lane);
    }
    const long long qi = (long long)blockIdx.x * kWave + lane;
    const bool valid = qi < nq;
    double px = 0, py = 0, pz = 0;
    if (valid) {
        px = qx[qi];
        py = qy[qi];
        pz = qz[qi];
    }
    double qlo[3], qhi[3];
    qlo[0] = wave_min_d(valid ? px : DBL_MAX);
    qlo[1] = wave_min_d(valid ? py : DBL_MAX);
    qlo[2] = wave_min_d(valid ? pz : DBL_MAX);
    qhi[0] = wave_max_d(valid ? px : -DBL_MAX);
    qhi[1] = wave_max_d(valid ? py : -DBL_MAX);
    qhi[2] = wave_max_d(valid ? pz : -DBL_MAX);

    double td[kTop];
    unsigned ti[kTop];
#pragma unroll
    for (int j = 0; j < kTop; j++) {
        td[j] = __longlong_as_double(0x7ff0000000000000ll);  // +inf
        ti[j] = 0xffffffffu;
    }
    Key kth;
    kth.d = td[0];
    kth.i = ti[0];
    int cnt = 0;
    double R2 = __longlong_as_double(0x7ff0000000000000ll);

    int sp = 0;
    if (lane == 0) s_stack[0] = ((lv.nlev - 1) << 26) | 0;
    sp = 1;
    __syncthreads();
    while (sp > 0) {
        int code = s_stack[sp - 1];
        sp--;
        int lvl = code >> 26;
        long long j = code & ((1 << 26) - 1);
        Box bx = boxes[s_off[lvl] + j];
        if (box_gap2(bx, qlo, qhi) > R2) continue;
        if (lvl > 0) {
            long long c0 = 2 * j, c1 = 2 * j + 1;
            __syncthreads();
            if (c1 < s_n[lvl - 1]) {
                Box b0 = boxes[s_off[lvl - 1] + c0];
                Box b1 = boxes[s_off[lvl - 1] + c1];
                double g0 = box_gap2(b0, qlo, qhi), g1 = box_gap2(b1, qlo, qhi);
                long long nearc = g0 <= g1 ? c0 : c1, farc = g0 <= g1 ? c1 : c0;
                if (lane == 0) {
                    s_stack[sp] = ((lvl - 1) << 26) | (int)farc;
                    s_stack[sp + 1] = ((lvl - 1) << 26) | (int)nearc;
                }
                sp += 2;
            } else {
                if (lane == 0) s_stack[sp] = ((lvl - 1) << 26) | (int)c0;
                sp += 1;
            }
            __syncthreads();
            continue;
        }
        // leaf bucket j: does any lane need it?
        double lq[3] = {px, py, pz};
        double lg0 = gap1(bx.lo[0], bx.hi[0], lq[0], lq[0]);
        double lg1 = gap1(bx.lo[1], bx.hi[1], lq[1], lq[1]);
        double lg2 = gap1(bx.lo[2], bx.hi[2], lq[2], lq[2]);
        double lgap = (lg0 * lg0 + lg1 * lg1) + lg2 * lg2;
        bool need = valid && !(lgap > kth.d);
        if (__ballot(need) == 0ull) continue;
        long long base = j * kWave;
        int ncand = (int)min((long long)kWave, nc - base);
        __syncthreads();
        if (lane < ncand) {
            s_x[lane] = cx[base + lane];
            s_y[lane] = cy[base + lane];
            s_z[lane] = cz[base + lane];
            s_id[lane] = cid[base + lane];
        }
        __syncthreads();
        for (int c0 = 0; c0 < ncand; c0 += kChunkC) {
            int c1 = min(c0 + kChunkC, ncand);
            for (int c = c0; c < c1; c++) {
                Key key;
                key.d = d2_ref(px, py, pz, s_x[c], s_y[c], s_z[c]);
                key.i = s_id[c];
                if (need && key_lt(key, kth)) {
                    s_bd[cnt][lane] = key.d;
                    s_bi[cnt][lane] = key.i;
                    cnt++;
                }
            }
            if (__ballot(cnt > kBuf - kChunkC) != 0ull) {
                flush_buffer(td, ti, cnt, s_bd, s_bi, lane);
                cnt = 0;
                kth = kth_of(td, ti, k);
            }
        }
        R2 = wave_max_d(valid ? kth.d : 0.0);
    }
    // final merge
    if (__ballot(cnt > 0) != 0ull) {
        flush_buffer(td, ti, cnt, s_bd, s_bi, lane);
        kth = kth_of(td, ti, k);
    }
    if (!valid) return;
    const long long q = qid[qi];
    if (nbr) {
        long long *row = nbr + q * k;
#pragma unroll
        for (int t = 0; t < kTop; t++)
            if (t < k) row[t] = (long long)ti[t];
    }
    if (rho) {
        float neardist = sqrtf((float)kth.d);
        float h = (float)((double)neardist / 2.0);
        float s = 0.0f;
#pragma unroll
        for (int t = 0; t < kTop; t++) {
            if (t < k) {
                float r = sqrtf((float)td[t]);
                float w = W4(r, h);
                if (rho_mode == 2) w = w * (float)mass[ti[t]];
                s += w;
            }
        }
        rho[q] = s;
    }
}

inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

template <class P>
hipError_t knn(const P *x, const P *y, const P *z, const P *m, long long np, const int *type,
               int cand_ptype, int k, long long *nbr, float *rho, int rho_mode, Workspace &ws,
               hipStream_t s) {
    if (k < 1 || k > kTop || np < 1) return hipErrorInvalidValue;
    timing_mark("knn_prep", s, false);
    const bool subset = type != nullptr && cand_ptype != PGALF_TYPE_ALL;
    // bounding box of every particle (queries include all candidates)
    unsigned long long *mm = ws.alloc_n<unsigned long long>(6);
    if (!mm) return hipErrorOutOfMemory;
    PGALF_HIP_TRY(hipMemsetAsync(mm, 0xff, 3 * sizeof(unsigned long long), s));
    PGALF_HIP_TRY(hipMemsetAsync(mm + 3, 0x00, 3 * sizeof(unsigned long long), s));
    bbox_k<P><<<1024, 256, 0, s>>>(x, y, z, np, mm);

    // queries: all particles, Morton order
    unsigned *qk = ws.alloc_n<unsigned>(np), *qv = ws.alloc_n<unsigned>(np);
    unsigned *qk2 = ws.alloc_n<unsigned>(np), *qv2 = ws.alloc_n<unsigned>(np);
    if (!qk || !qv || !qk2 || !qv2) return hipErrorOutOfMemory;
    morton_k<P><<<nblk(np, 256), 256, 0, s>>>(x, y, z, nullptr, np, mm, qk, qv);
    PGALF_HIP_TRY(radix_sort_pairs(&qk, &qv, qk2, qv2, np, 30, ws, s));
    double *qx = ws.alloc_n<double>(np), *qy = ws.alloc_n<double>(np),
           *qz = ws.alloc_n<double>(np);
    if (!qx || !qy || !qz) return hipErrorOutOfMemory;
    gather_sorted_k<P><<<nblk(np, 256), 256, 0, s>>>(x, y, z, qv, np, qx, qy, qz);

    long long nc = np;
    const double *cx = qx, *cy = qy, *cz = qz;
    const unsigned *cid = qv;
    if (subset) {
        unsigned *flags = ws.alloc_n<unsigned>(np), *pos = ws.alloc_n<unsigned>(np);
        unsigned *tot = ws.alloc_n<unsigned>(1);
        unsigned *sel = ws.alloc_n<unsigned>(np);
        if (!flags || !pos || !tot || !sel) return hipErrorOutOfMemory;
        select_flags_k<<<nblk(np, 256), 256, 0, s>>>(type, np, cand_ptype, flags);
        PGALF_HIP_TRY(exclusive_scan_u32(flags, pos, np, tot, ws, s));
        select_compact_k<<<nblk(np, 256), 256, 0, s>>>(flags, pos, np, sel);
        unsigned h_tot = 0;
        PGALF_HIP_TRY(hipMemcpyAsync(&h_tot, tot, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        PGALF_HIP_TRY(hipStreamSynchronize(s));
        nc = h_tot;
        if (nc < k) return hipErrorNotReady;  // mapped to PGALF_ERR_SHORT_LIST
        unsigned *ck = ws.alloc_n<unsigned>(nc), *cv = ws.alloc_n<unsigned>(nc);
        unsigned *ck2 = ws.alloc_n<unsigned>(nc), *cv2 = ws.alloc_n<unsigned>(nc);
        double *sx = ws.alloc_n<double>(nc), *sy = ws.alloc_n<double>(nc),
               *sz = ws.alloc_n<double>(nc);
        if (!ck || !cv || !ck2 || !cv2 || !sx || !sy || !sz) return hipErrorOutOfMemory;
        morton_k<P><<<nblk(nc, 256), 256, 0, s>>>(x, y, z, sel, nc, mm, ck, cv);
        PGALF_HIP_TRY(radix_sort_pairs(&ck, &cv, ck2, cv2, nc, 30, ws, s));
        gather_sorted_k<P><<<nblk(nc, 256), 256, 0, s>>>(x, y, z, cv, nc, sx, sy, sz);
        cx = sx;
        cy = sy;
        cz = sz;
        cid = cv;
    } else if (np < k) {
        return hipErrorNotReady;
    }

    // box tree over candidate buckets
    Levels lv;
    long long nb = (nc + kWave - 1) / kWave;
    if (nb >= (1LL << 26)) return hipErrorInvalidValue;
    lv.nb = nb;
    lv.nlev = 1;
    while (level_count(nb, lv.nlev - 1) > 1) lv.nlev++;
    long long total = level_offset(nb, lv.nlev);
    Box *boxes = ws.alloc_n<Box>((size_t)total);
    if (!boxes) return hipErrorOutOfMemory;
    leaf_boxes_k<<<(unsigned)nb, kWave, 0, s>>>(cx, cy, cz, nc, boxes);
    for (int l = 1; l < lv.nlev; l++) {
        long long nl = level_count(nb, l);
        parent_boxes_k<<<nblk(nl, 256), 256, 0, s>>>(boxes + level_offset(nb, l - 1),
                                                     level_count(nb, l - 1),
                                                     boxes + level_offset(nb, l), nl);
    }
    timing_mark("knn_prep", s, true);
    timing_mark("knn_query", s, false);
    long long nqb = (np + kWave - 1) / kWave;
    knn_query_k<P><<<(unsigned)nqb, kWave, 0, s>>>(qx, qy, qz, qv, np, cx, cy, cz, cid, nc,
                                                   boxes, lv, k, nbr, rho, rho_mode, m);
    timing_mark("knn_query", s, true);
    return hipGetLastError();
}

template hipError_t knn<float>(const float *, const float *, const float *, const float *,
                               long long, const int *, int, int, long long *, float *, int,
                               Workspace &, hipStream_t);
template hipError_t knn<double>(const double *, const double *, const double *,
                                const double *, long long, const int *, int, int, long long *,
                                float *, int, Workspace &, hipStream_t);

}  // namespace pgalf
