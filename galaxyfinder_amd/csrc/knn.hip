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
//   * queries and candidates are sorted by their 30-bit Hilbert index (10
//     bits per axis over the bounding cube; stable radix sort, ties by
//     index).  Hilbert order has no jumps, so 8 consecutive particles form a
//     compact bucket;
//   * candidate buckets of 8 get FP64 bounding boxes and an implicit 8-ary
//     tree of boxes (8 consecutive nodes per parent);
//   * a wavefront holds eight independent query groups of 8 lanes (8
//     consecutive sorted queries -- one bucket -- one per lane).  Each group
//     fills every lane's list from the 4 buckets around its own (a 32-key
//     odd-even merge sort) and merges the 8 buckets around them (4 on
//     either side) straight in, 16 keys at a time (about half of them would pass a bound
//     test, so testing and buffering them costs more), which gives every
//     lane a k-th distance.  It then walks the tree depth first, nearest
//     child first: a round pops one node and its 8 lanes test the 8 children
//     against the group's box and the union of the lanes' current balls.
//     Passing buckets go to a list, ordered by distance class (the first
//     list's 4 nearest buckets are merged straight in, like the seed) and
//     scanned in batches: each batch takes the first two buckets of an 8-entry window
//     that some lane's ball still reaches, stages their candidates in LDS,
//     and every lane tests them against its own k-th key.  Accepted
//     candidates go to a per-lane LDS buffer; when any lane's buffer may
//     overflow, all lanes merge its first 16 keys into their sorted top-32
//     list held in registers (odd-even sort of 16 + bitonic merge) and move
//     the rest down; at the end of a list every key is merged;
//   * keys are (FP64 dist2, uint32 index) compared lexicographically, so the
//     lists are exact and deterministic regardless of visiting order.
// Box pruning is conservative: box gaps are formed with the same monotone
// FP64 operations as the distances, from box faces that are member
// coordinates, so a gap never exceeds a member's computed distance.
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <vector>

#include "pgalf_internal.h"

namespace pgalf {
namespace {

constexpr int kWave = 64;
constexpr int kTop = 32;     // list capacity (k <= 32, MAX_NUM_NEAR tree.h:20)
#ifndef PGALF_KNN_MERGE
#define PGALF_KNN_MERGE 16
#endif
constexpr int kBuf = PGALF_KNN_MERGE;  // merge chunk (sorted, then merged into the top-32)
// Capacities below can be set at compile time (-DPGALF_KNN_...): they trade
// LDS per wavefront (occupancy) against merge and scan frequency.
#ifndef PGALF_KNN_BUFCAP
#define PGALF_KNN_BUFCAP 32
#endif
#ifndef PGALF_KNN_LISTCAP
#define PGALF_KNN_LISTCAP 64
#endif
#ifndef PGALF_KNN_BATCH
#define PGALF_KNN_BATCH 2
#endif
constexpr int kBufCap = PGALF_KNN_BUFCAP;  // per-lane LDS append buffer (sorted-candidate indices)
#ifndef PGALF_KNN_CHUNK
#define PGALF_KNN_CHUNK 8
#endif
constexpr int kChunkC = PGALF_KNN_CHUNK;  // candidates between overflow checks
constexpr int kMaxLevels = 24;
constexpr int kB = 8;        // candidates per bucket
constexpr int kG = 8;        // lanes per query group (one candidate of a bucket each)
constexpr int kGroups = kWave / kG;
constexpr unsigned long long kGMask = (1ull << kG) - 1ull;
constexpr int kFan = 8;      // box-tree fan-out
constexpr int kPar = kG / kFan;  // parents expanded per walk round
#ifndef PGALF_KNN_SEEDSPAN
#define PGALF_KNN_SEEDSPAN 12
#endif
#ifndef PGALF_KNN_FIRSTSCAN
#define PGALF_KNN_FIRSTSCAN 16
#endif
constexpr int kSeedFill = kTop / kB;  // buckets of the direct seed fill
constexpr int kSeedSpan = PGALF_KNN_SEEDSPAN;  // seed buckets in all (a window around the fill)
#ifndef PGALF_KNN_DIRECT
#define PGALF_KNN_DIRECT 8
#endif
constexpr int kDirect = PGALF_KNN_DIRECT;  // window buckets merged without a bound test
#ifndef PGALF_KNN_WALK_DIRECT
#define PGALF_KNN_WALK_DIRECT 4
#endif
constexpr int kWalkDirect = PGALF_KNN_WALK_DIRECT;  // first-list buckets merged the same way
#ifndef PGALF_KNN_FLUSH_ONE
#define PGALF_KNN_FLUSH_ONE 1
#endif
constexpr bool kFlushOne = PGALF_KNN_FLUSH_ONE != 0;  // mid-scan flushes merge one chunk
constexpr int kFCap = PGALF_KNN_LISTCAP;  // bucket-list capacity per group
constexpr int kStackCap = 96;  // depth-first walk stack per group
constexpr int kFirstScan = PGALF_KNN_FIRSTSCAN;  // the first list is scanned early to tighten the bounds
constexpr int kLevelShift = 5;  // stack entry = node << 5 | level
constexpr int kBatch = PGALF_KNN_BATCH;  // buckets staged in LDS per scan batch
constexpr int kStage = kBatch * kB;  // staged candidates per group
constexpr int kStageStride = 2 * kStage + 1;  // 16-byte units per group: groups hit other banks
#ifndef PGALF_KNN_RINGS
#define PGALF_KNN_RINGS 4
#endif
#ifndef PGALF_KNN_RING_STEPS
#define PGALF_KNN_RING_STEPS 2
#endif
constexpr int kRings = PGALF_KNN_RINGS;  // distance classes of the near-first bucket order
constexpr int kRingSteps = PGALF_KNN_RING_STEPS;  // classes per R (the group's largest ball)
constexpr int kRingShift = 28;  // list entry = bucket | ring << 28
constexpr int kBucketMask = (1 << kRingShift) - 1;
static_assert(kG == kB, "a scan batch loads one candidate of each bucket per group lane");
static_assert(kG % kFan == 0 && kPar >= 1, "a walk round expands whole parents");
static_assert(kTop % kB == 0, "the direct seed fill takes whole buckets");
static_assert(kPar == 1, "the depth-first walk expands one parent per round");
static_assert(kBufCap % kBuf == 0 && kBufCap % kChunkC == 0 && kBufCap >= 24,
              "buffer holds whole merge chunks");
static_assert(kBatch >= 1 && kBatch <= kG, "one group lane tests each bucket of a batch");
static_assert(kFCap >= 32 + 8 && kFCap % 8 == 0, "list must take the first scan and a round");
static_assert(kSeedSpan >= kSeedFill && kFirstScan + kG <= kFCap, "seed window and first scan");
constexpr int kDM = 2 * kB;  // keys per direct-merge chunk (two buckets)
static_assert(kDirect % 2 == 0 && kDirect <= kSeedSpan - kSeedFill,
              "direct merges take two buckets per chunk");
static_assert(kWalkDirect % 2 == 0 && kWalkDirect <= kFirstScan,
              "walk direct merges take two buckets per chunk, from the first list");
// The walk keeps at most kFan entries of the deepest level and kFan-1 of every
// level above it on the stack.
constexpr int stack_need(int nlev) { return kFan + (kFan - 1) * (nlev > 3 ? nlev - 3 : 0); }
constexpr int kMaxWalkLevels = 14;
static_assert(stack_need(kMaxWalkLevels) <= kStackCap, "walk stack too small");

struct Box {
    double lo[3], hi[3];
};

// Workgroups go to the 8 XCDs round-robin (block b to XCD b % 8), and each
// XCD has its own L2.  Neighbouring query waves read mostly the same
// candidates, so each XCD is given runs of `chunk` consecutive logical waves
// (the schedule's knn_xcd_chunk), the runs dealt to the XCDs in turn
// (contiguous ranges per XCD would unbalance them where wave costs vary
// along the order, e.g. a batch sorted by halo size).  Block b runs wave
// xcd_block(b, n, chunk), a bijection of [0, n): blocks past the last whole
// round of runs keep their own index.
constexpr unsigned kXcds = 8;
__host__ __device__ inline unsigned xcd_block(unsigned b, unsigned n, unsigned chunk) {
    if (chunk == 0) return b;
    const unsigned round = kXcds * chunk;
    if (b >= n / round * round) return b;
    const unsigned x = b % kXcds, i = b / kXcds;  // the i-th block of XCD x
    return ((i / chunk) * kXcds + x) * chunk + i % chunk;
}

// Implicit 8-ary box tree over nb leaf buckets: level l holds
// ceil(nb / 8^l) nodes.  One tree per halo; the box of node j of level l is
// boxes[loff[l] + j].
struct KHalo {
    long long c0;        // first candidate slot (a multiple of kB)
    long long nc;        // candidates
    long long q0, nq;    // query slots [q0, q0 + nq)
    long long idx_base;  // subtracted from candidate ids in the output lists
    long long nb;        // leaf buckets
    int nlev;
    int kb;              // slot bits of the packed keys: nb * kB <= 2^kb
    long long loff[kMaxWalkLevels];
};
// Largest slot-bit count of the packed-key query (keys keep 51 - kb bits of
// the distance); larger halos take the exact query only.
constexpr int kMaxKeyBits = 30;
inline int key_bits(long long nslots) {
    int b = 1;
    while ((1LL << b) < nslots) b++;
    return b;
}
// One wavefront of queries: kWave consecutive query slots of one halo.
struct KWave {
    int halo;
    long long q0;
};

__host__ __device__ inline long long level_count(long long nb, int l) {
    long long d = 1LL << (3 * l);
    return (nb + d - 1) / d;
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

// Compare-exchange of two keys (a <= b afterwards).  The distances are
// ordered by min/max (equal distances are interchangeable, and list keys are
// never NaN); only the indices need the lexicographic select.
__device__ inline void cas2(double &ad, unsigned &ai, double &bd, unsigned &bi) {
    const bool sw = klt(bd, bi, ad, ai);
    const double lo = fmin(ad, bd), hi = fmax(ad, bd);
    const unsigned t1 = ai, t3 = bi;
    ad = lo;
    bd = hi;
    ai = sw ? t3 : t1;
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

// Bounding box: per-block reduction, then one atomic per block and value.
template <class P>
__global__ void __launch_bounds__(256) bbox_k(const P *x, const P *y, const P *z, long long np,
                                              unsigned long long *mm /* 3 min, 3 max */) {
    __shared__ double s_v[6][4];
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
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = wave_min_d(lo[a]);
        hi[a] = wave_max_d(hi[a]);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            s_v[a][w] = lo[a];
            s_v[3 + a][w] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int a = threadIdx.x;
        double v = s_v[a][0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++)
            v = a < 3 ? fmin(v, s_v[a][i]) : fmax(v, s_v[a][i]);
        if (a < 3) atomicMin(&mm[a], ord_d(v));
        else atomicMax(&mm[a], ord_d(v));
    }
}

// 3-D Hilbert index of 10-bit coordinates (Skilling's axes-to-transpose,
// then bit interleave with x most significant).
__device__ inline unsigned hilbert3(unsigned x, unsigned y, unsigned z) {
    unsigned X[3] = {x, y, z};
#pragma unroll
    for (unsigned Q = 1u << 9; Q > 1; Q >>= 1) {
        const unsigned P = Q - 1;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                unsigned t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    unsigned t = 0;
#pragma unroll
    for (unsigned Q = 1u << 9; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    unsigned key = 0;
#pragma unroll
    for (int b = 9; b >= 0; b--)
        key = (key << 3) | (((X[0] >> b) & 1u) << 2) | (((X[1] >> b) & 1u) << 1) |
              ((X[2] >> b) & 1u);
    return key;
}

// Hilbert key of point v on the 1024^3 lattice over the cube of the box mm
// (3 ordered minima, 3 ordered maxima).
__device__ inline unsigned hilbert_of(const double v[3], const unsigned long long *mm) {
    double lo[3], ext = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = unord_d(mm[a]);
        ext = fmax(ext, unord_d(mm[3 + a]) - lo[a]);
    }
    double sc = ext > 0.0 ? 1024.0 / ext : 0.0;
    unsigned q[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double t = (v[a] - lo[a]) * sc;
        int qi = (int)t;
        q[a] = (unsigned)(qi < 0 ? 0 : (qi > 1023 ? 1023 : qi));
    }
    return hilbert3(q[0], q[1], q[2]);
}

// Hilbert keys of particles idx[0..n) (idx == nullptr: identity).
template <class P>
__global__ void hilbert_k(const P *x, const P *y, const P *z, const unsigned *idx, long long n,
                          const unsigned long long *mm, unsigned *keys, unsigned *vals,
                          P *packed) {
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    unsigned i = idx ? idx[s] : (unsigned)s;
    double v[3] = {(double)x[i], (double)y[i], (double)z[i]};
    keys[s] = hilbert_of(v, mm);
    vals[s] = i;
    if (packed) {  // (x, y, z, 0) per particle, for one-load gathers after the sort
        P *d = packed + 4 * (size_t)i;
        d[0] = x[i];
        d[1] = y[i];
        d[2] = z[i];
        d[3] = P(0);
    }
}

// Sorted FP64 copies from the packed (x, y, z, 0) records: one 16-byte load
// per float particle (two for double).
template <class P>
__global__ void gather_packed_k(const P *packed, const unsigned *perm, long long n, double *sx,
                                double *sy, double *sz) {
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const size_t i = perm[s];
    if constexpr (sizeof(P) == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(packed + 4 * i);
        sx[s] = (double)v.x;
        sy[s] = (double)v.y;
        sz[s] = (double)v.z;
    } else {
        const double2 a = *reinterpret_cast<const double2 *>(packed + 4 * i);
        const double2 b = *reinterpret_cast<const double2 *>(packed + 4 * i + 2);
        sx[s] = (double)a.x;
        sy[s] = (double)a.y;
        sz[s] = (double)b.x;
    }
}

// ---- batched preparation (one block per particle chunk of one halo) -------
template <class P>
__global__ void __launch_bounds__(256) bbox_batch_k(const P *x, const P *y, const P *z,
                                                    const PChunk *chunks,
                                                    unsigned long long *mm /* 6 per halo */) {
    __shared__ double s_v[6][4];
    const PChunk c = chunks[blockIdx.x];
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (long long i = c.b0 + threadIdx.x; i < c.b1; i += blockDim.x) {
        double v[3] = {(double)x[i], (double)y[i], (double)z[i]};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = fmin(lo[a], v[a]);
            hi[a] = fmax(hi[a], v[a]);
        }
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = wave_min_d(lo[a]);
        hi[a] = wave_max_d(hi[a]);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            s_v[a][w] = lo[a];
            s_v[3 + a][w] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int a = threadIdx.x;
        double v = s_v[a][0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++)
            v = a < 3 ? fmin(v, s_v[a][i]) : fmax(v, s_v[a][i]);
        unsigned long long *hm = mm + 6 * (long long)c.halo;
        if (a < 3) atomicMin(&hm[a], ord_d(v));
        else atomicMax(&hm[a], ord_d(v));
    }
}

// Minima start at the largest ordered value, maxima at the smallest.
__global__ void bbox_init_k(unsigned long long *mm, int nhalo) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 6LL * nhalo) mm[i] = (i % 6) < 3 ? ~0ull : 0ull;
}

// Hilbert key of every particle within its halo's own box, and its halo.
template <class P>
__global__ void __launch_bounds__(256) hilbert_batch_k(const P *x, const P *y, const P *z,
                                                       const PChunk *chunks,
                                                       const unsigned long long *mm,
                                                       unsigned *keys, unsigned *vals,
                                                       unsigned *halo_of) {
    const PChunk c = chunks[blockIdx.x];
    const unsigned long long *hm = mm + 6 * (long long)c.halo;
    for (long long i = c.b0 + threadIdx.x; i < c.b1; i += blockDim.x) {
        double v[3] = {(double)x[i], (double)y[i], (double)z[i]};
        keys[i] = hilbert_of(v, hm);
        vals[i] = (unsigned)i;
        halo_of[i] = (unsigned)c.halo;
    }
}

__global__ void halo_keys_k(const unsigned *vals, const unsigned *halo_of, long long n,
                            unsigned *hk) {
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) hk[s] = halo_of[vals[s]];
}

// (halo, Hilbert)-sorted particles into the padded candidate layout: halo
// h's s-th particle goes to slot c0 + s; the pad slots up to the next
// multiple of kB hold NaN coordinates and id 0xffffffff.
template <class P>
__global__ void gather_padded_k(const P *x, const P *y, const P *z, const unsigned *perm,
                                const unsigned *hsorted, const KHalo *halos, long long n,
                                double *sx, double *sy, double *sz, unsigned *sid) {
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const KHalo &hd = halos[hsorted[s]];
    const long long p = hd.c0 + (s - hd.idx_base);
    const unsigned i = perm[s];
    sx[p] = (double)x[i];
    sy[p] = (double)y[i];
    sz[p] = (double)z[i];
    sid[p] = i;
}

__global__ void pad_slots_k(const KHalo *halos, int nhalo, double *sx, double *sy, double *sz,
                            unsigned *sid) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nhalo * kB) return;
    const KHalo &hd = halos[t / kB];
    const long long j = hd.nc + t % kB;
    if (j >= hd.nb * kB) return;
    const double nan = __longlong_as_double(0x7ff8000000000000ll);
    sx[hd.c0 + j] = nan;
    sy[hd.c0 + j] = nan;
    sz[hd.c0 + j] = nan;
    sid[hd.c0 + j] = 0xffffffffu;
}

// Internal nodes [b0, b1) of one level: union of their children's boxes.
__global__ void node_boxes_k(Box *boxes, long long b0, long long b1, long long nleaf,
                             const long long *child_begin, const unsigned char *child_cnt) {
    long long b = b0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= b1) return;
    const long long t = b - nleaf;
    const long long c0 = child_begin[t];
    Box a = boxes[c0];
    for (long long c = c0 + 1; c < c0 + child_cnt[t]; c++) {
        Box q = boxes[c];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            a.lo[d] = fmin(a.lo[d], q.lo[d]);
            a.hi[d] = fmax(a.hi[d], q.hi[d]);
        }
    }
    boxes[b] = a;
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

// One lane per candidate slot (coalesced loads); the kB lanes of a bucket
// reduce its box by shuffles.  fmin/fmax skip NaN pad slots.  Launch with a
// multiple of kB threads per block covering nb * kB slots.
__global__ void leaf_boxes_k(const double *sx, const double *sy, const double *sz, long long n,
                             long long nb, Box *boxes) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    if (i < n) {
        const double v[3] = {sx[i], sy[i], sz[i]};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = fmin(lo[a], v[a]);
            hi[a] = fmax(hi[a], v[a]);
        }
    }
#pragma unroll
    for (int o = 1; o < kB; o <<= 1)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = fmin(lo[a], __shfl_xor(lo[a], o, kB));
            hi[a] = fmax(hi[a], __shfl_xor(hi[a], o, kB));
        }
    const long long b = i / kB;
    if (i % kB == 0 && b < nb) {
        Box bx;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            bx.lo[a] = lo[a];
            bx.hi[a] = hi[a];
        }
        boxes[b] = bx;
    }
}

// gather_packed_k and leaf_boxes_k in one pass, for a tree whose candidates
// are all the sorted queries: thread s gathers sorted particle s and the kB
// lanes of a bucket reduce its box as leaf_boxes_k does (threads past n
// hold empty extents).
template <class P>
__global__ void gather_leaf_k(const P *packed, const unsigned *perm, long long n, double *sx,
                              double *sy, double *sz, long long nb, Box *boxes) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    if (s < n) {
        const size_t i = perm[s];
        double v[3];
        if constexpr (sizeof(P) == 4) {
            const float4 f = *reinterpret_cast<const float4 *>(packed + 4 * i);
            v[0] = (double)f.x;
            v[1] = (double)f.y;
            v[2] = (double)f.z;
        } else {
            const double2 a = *reinterpret_cast<const double2 *>(packed + 4 * i);
            const double2 b = *reinterpret_cast<const double2 *>(packed + 4 * i + 2);
            v[0] = (double)a.x;
            v[1] = (double)a.y;
            v[2] = (double)b.x;
        }
        sx[s] = v[0];
        sy[s] = v[1];
        sz[s] = v[2];
#pragma unroll
        for (int a = 0; a < 3; a++) lo[a] = hi[a] = v[a];
    }
#pragma unroll
    for (int o = 1; o < kB; o <<= 1)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = fmin(lo[a], __shfl_xor(lo[a], o, kB));
            hi[a] = fmax(hi[a], __shfl_xor(hi[a], o, kB));
        }
    const long long b = s / kB;
    if (s % kB == 0 && b < nb) {
        Box bx;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            bx.lo[a] = lo[a];
            bx.hi[a] = hi[a];
        }
        boxes[b] = bx;
    }
}

__global__ void parent_boxes_k(const Box *child, long long nchild, Box *parent,
                               long long nparent) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nparent) return;
    Box a = child[kFan * j];
    for (long long c = kFan * j + 1; c < min(nchild, kFan * (j + 1)); c++) {
        Box b = child[c];
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

__device__ inline double grp_min_d(double v) {
#pragma unroll
    for (int o = kG / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double grp_max_d(double v) {
#pragma unroll
    for (int o = kG / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
#ifdef PGALF_KNN_STATS
__device__ inline int grp_max_i(int v) {
#pragma unroll
    for (int o = kG / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
#endif
__device__ inline int wave_max_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Debug-only per-wave counters (PGALF_DEBUG): events and shader-clock cycles.
#ifdef PGALF_KNN_STATS
constexpr int kDbgN = 64;
#else
constexpr int kDbgN = 16;
#endif
__device__ inline unsigned long long dbg_clock(const unsigned long long *dbg) {
    return dbg ? (unsigned long long)clock64() : 0ull;
}
__device__ inline void dbg_cycles(unsigned long long *dbg, int i, unsigned long long t0, int lane) {
    if (dbg && lane == 0) atomicAdd(&dbg[i], (unsigned long long)clock64() - t0);
}

// Sorting networks written as compile-time recursions over split (dist,
// index) arrays so every index is a constant and the lists stay in
// registers.
//
// Batcher's odd-even merge sort (1968): 63
// compare-exchanges for 16 keys and 191 for 32, against the bitonic sort's 80
// and 240.  Keys are totally ordered (distinct indices), so every sorting
// network gives the same output.
template <int N, int LO, int HI, int R>
__device__ inline void oe_merge(double (&d)[N], unsigned (&x)[N]) {
    constexpr int STEP = 2 * R;
    if constexpr (STEP < HI - LO) {
        oe_merge<N, LO, HI, STEP>(d, x);
        oe_merge<N, LO + R, HI, STEP>(d, x);
#pragma unroll
        for (int i = LO + R; i < HI - R; i += STEP) cas2(d[i], x[i], d[i + R], x[i + R]);
    } else {
        cas2(d[LO], x[LO], d[LO + R], x[LO + R]);
    }
}
template <int N, int LO = 0, int HI = N - 1>
__device__ inline void oe_sort(double (&d)[N], unsigned (&x)[N]) {
    if constexpr (HI - LO >= 1) {
        constexpr int MID = LO + (HI - LO) / 2;
        oe_sort<N, LO, MID>(d, x);
        oe_sort<N, MID + 1, HI>(d, x);
        oe_merge<N, LO, HI, 1>(d, x);
    }
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

// Merge a chunk of N keys into the sorted top-32: sort the chunk, take the
// elementwise minimum of the top against the chunk reversed and padded with
// +inf to 32 -- the 32 smallest, as a bitonic sequence -- and clean.
template <int N>
__device__ inline void merge_top(double (&td)[kTop], unsigned (&ti)[kTop], double (&bd)[N],
                                 unsigned (&bi)[N]) {
    oe_sort<N>(bd, bi);
#pragma unroll
    for (int j = 0; j < N; j++) {
        double cd = bd[N - 1 - j];
        unsigned ci = bi[N - 1 - j];
        constexpr int o = kTop - N;  // reversed buffer meets the top's tail
        bool take = klt(cd, ci, td[o + j], ti[o + j]);
        td[o + j] = take ? cd : td[o + j];
        ti[o + j] = take ? ci : ti[o + j];
    }
    bitonic_clean<kTop, kTop / 2>(td, ti);
}

// k-th smallest of the sorted list as a masked maximum (a select on
// t == k-1 would be folded into a dynamic index).
__device__ inline Key kth_of(const double (&td)[kTop], const unsigned (&ti)[kTop], int k) {
    Key r;
    if (k == kTop) {  // the common case: the last slot
        r.d = td[kTop - 1];
        r.i = ti[kTop - 1];
        return r;
    }
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

// Merge a lane's buffer of cnt sorted-candidate indices into its top-32:
// keys are recomputed from the candidate arrays (same arithmetic as the
// scan, so the same FP64 values), then merged kBuf at a time.  Wave uniform.
__device__ inline void flush_buffer(double (&td)[kTop], unsigned (&ti)[kTop], int cnt,
                                    const unsigned (*s_bc)[kWave], int lane, double px,
                                    double py, double pz, const double *cx, const double *cy,
                                    const double *cz, const unsigned *cid,
                                    int nch = kBufCap / kBuf) {
#pragma unroll 1
    for (int ch = 0; ch < nch; ch++) {  // not unrolled: bounds live loads
        if (__ballot(cnt > ch * kBuf) == 0ull) break;
        double bd[kBuf];
        unsigned bi[kBuf];
#pragma unroll
        for (int t = 0; t < kBuf; t++) {
            const int e = ch * kBuf + t;
            bd[t] = __longlong_as_double(0x7ff0000000000000ll);
            bi[t] = 0xffffffffu;
            if (e < cnt) {
                unsigned sidx = s_bc[e][lane];
                bd[t] = d2_ref(px, py, pz, cx[sidx], cy[sidx], cz[sidx]);
                bi[t] = cid[sidx];
            }
        }
        merge_top(td, ti, bd, bi);
    }
}

// Per-lane query state of the kNN kernel.
struct QState {
    double td[kTop];
    unsigned ti[kTop];
    Key kth;
    int cnt;
#ifdef PGALF_KNN_STATS
    int acc;  // candidates accepted into the buffer so far
#endif
};

// LDS staging area of one scan batch: kBatch buckets per group, one 32-byte
// record per candidate in two 16-byte units -- {x, y} and {z, sorted index,
// particle index} -- read with two 128-bit loads.  An empty slot has a NaN x,
// so its distance never compares below a bound.  Between batches the area
// doubles as the ring sort's scratch (kGroups * kFCap ints).
struct Stage {
    uint4 u[kGroups * kStageStride];
};
__device__ inline void stage_put(Stage &sg, int grp, int slot, double x, double y, double z,
                                 unsigned e, unsigned id) {
    const int b = grp * kStageStride + 2 * slot;
    const unsigned long long bx = (unsigned long long)__double_as_longlong(x);
    const unsigned long long by = (unsigned long long)__double_as_longlong(y);
    const unsigned long long bz = (unsigned long long)__double_as_longlong(z);
    sg.u[b] = make_uint4((unsigned)bx, (unsigned)(bx >> 32), (unsigned)by, (unsigned)(by >> 32));
    sg.u[b + 1] = make_uint4((unsigned)bz, (unsigned)(bz >> 32), e, id);
}
__device__ inline double hilo_d(unsigned lo, unsigned hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Does box b come within the current k-th distance of any query of group
// grp (points and k-th distances of the wave in LDS; invalid lanes hold -1)?
__device__ inline bool union_pass(const Box &b, int grp, const double *s_qx, const double *s_qy,
                                  const double *s_qz, const double *s_qk) {
    bool pass = false;
#pragma unroll 4
    for (int j = 0; j < kG; j++) {
        const int L = grp * kG + j;
        double x = s_qx[L], y = s_qy[L], z = s_qz[L];
        double g0 = gap1(b.lo[0], b.hi[0], x, x);
        double g1 = gap1(b.lo[1], b.hi[1], y, y);
        double g2 = gap1(b.lo[2], b.hi[2], z, z);
        double a = g0 * g0;
        double bb = g1 * g1;
        double c = g2 * g2;
        pass |= !(((a + bb) + c) > s_qk[L]);
    }
    return pass;
}

// Scan the buckets listed in list[0..count) of this lane's group (count is
// group-uniform).  kBatch buckets at a time are staged in LDS -- each lane
// of the group loads one candidate per bucket, so the global-memory latency
// is paid once per batch -- then every lane runs over the staged candidates
// and appends those closer than its k-th key to its buffer.  Wave uniform.
__device__ inline void scan_list(QState &st, int *list, int count, bool valid, int k,
                                 double px, double py, double pz, const double *cx,
                                 const double *cy, const double *cz, const unsigned *cid,
                                 long long nc, const Box *leaves, Stage &sg,
                                 unsigned (*s_bc)[kWave], double *s_qx, double *s_qy,
                                 double *s_qz, double *s_qk, int lane, int grp, int gl,
                                 unsigned long long *dbg) {
    const unsigned long long t_scan = dbg_clock(dbg);
    const double nan = __longlong_as_double(0x7ff8000000000000ll);
    int ipos = 0;  // the group's read position in its list
    const int nit = wave_max_i((count + kBatch - 1) / kBatch);  // a bound on the batches
    for (int it = 0; it < nit; it++) {
        if (__ballot(ipos < count) == 0ull) break;
        // the batch: the first kBatch entries of the window [ipos, ipos + kG)
        // that some lane's current ball reaches (group lane t tests entry
        // ipos + t); the window then advances past the last one taken, or
        // past the whole window when fewer pass -- entries that no ball
        // reaches any more are dropped before their candidates are loaded
        const int wend = ipos + kG < count ? ipos + kG : count;
        int ent = 0;
        bool pass = false;
        if (ipos + gl < count) {
            ent = list[ipos + gl];
            pass = union_pass(leaves[ent & kBucketMask], grp, s_qx, s_qy, s_qz, s_qk);
        }
        const unsigned mm = (unsigned)((__ballot(pass) >> (grp * kG)) & kGMask);
        const int npass = __popc(mm);
        // passing entries packed, in order, at the end of the window; the
        // batch takes the first kBatch of them and the next window starts
        // right after (every lane has read its entry before any writes)
        const int first = wend - npass;
        if (pass) list[first + __popc(mm & ((1u << gl) - 1u))] = ent;
        const int nsel = npass < kBatch ? npass : kBatch;
        if (ipos < count) ipos = first + nsel;
        double vx[kBatch], vy[kBatch], vz[kBatch];
        unsigned ve[kBatch], vid[kBatch];
#pragma unroll
        for (int t = 0; t < kBatch; t++) {
            long long e = t < nsel ? (long long)(list[first + t] & kBucketMask) * kB + gl : -1;
            if (e >= nc) e = -1;
            vx[t] = nan;
            vy[t] = 0.0;
            vz[t] = 0.0;
            ve[t] = 0xffffffffu;
            vid[t] = 0xffffffffu;
            if (e >= 0) {
                vx[t] = cx[e];
                vy[t] = cy[e];
                vz[t] = cz[e];
                ve[t] = (unsigned)e;
                vid[t] = cid[e];
            }
        }
        const int wsel = wave_max_i(nsel);
        if (dbg && gl == 0) {
            // staged buckets that some lane's ball reaches, and bucket slots scanned
            atomicAdd(&dbg[3], (unsigned long long)nsel);
            atomicAdd(&dbg[7], (unsigned long long)wsel);
        }
#pragma unroll
        for (int t = 0; t < kBatch; t++)
            stage_put(sg, grp, t * kB + gl, vx[t], vy[t], vz[t], ve[t], vid[t]);
        if (dbg && lane == 0) atomicAdd(&dbg[0], 1ull);  // scan batches (wave)
        __syncthreads();
        const int nslots = wsel * kB;
        const uint4 *su = sg.u + grp * kStageStride;
        for (int h = 0; h < nslots; h += kChunkC) {
            // branch-free: every candidate's index is written at the lane's
            // count, which advances only when the key is below the bound
            // (the count is at most kBufCap - kChunkC on entry)
#pragma unroll
            for (int c = h; c < h + kChunkC; c++) {
                const uint4 a = su[2 * c], b = su[2 * c + 1];
                const double d = d2_ref(px, py, pz, hilo_d(a.x, a.y), hilo_d(a.z, a.w),
                                        hilo_d(b.x, b.y));
                const bool acc = valid & klt(d, b.w, st.kth.d, st.kth.i);
                s_bc[st.cnt][lane] = b.z;
                st.cnt += acc ? 1 : 0;
            }
            if (__ballot(st.cnt > kBufCap - kChunkC) != 0ull) {
                if (dbg && lane == 0) atomicAdd(&dbg[1], 1ull);  // mid-scan merges
                if (dbg && valid) atomicAdd(&dbg[13], (unsigned long long)st.cnt);
#ifdef PGALF_KNN_STATS
                st.acc += valid ? st.cnt : 0;
#endif
                const int wm = dbg ? wave_max_i(st.cnt) : 0;
                if (dbg && lane == 0) atomicAdd(&dbg[14], (unsigned long long)((wm + kBuf - 1) / kBuf));
                const unsigned long long t_m = dbg_clock(dbg);
                if constexpr (kFlushOne) {
                    // one chunk for every lane; a lane's keys past it move down
                    flush_buffer(st.td, st.ti, st.cnt, s_bc, lane, px, py, pz, cx, cy, cz, cid, 1);
                    const int rest = st.cnt > kBuf ? st.cnt - kBuf : 0;
                    const int wrest = wave_max_i(rest);
                    for (int t = 0; t < wrest; t++)
                        if (t < rest) s_bc[t][lane] = s_bc[kBuf + t][lane];
                    st.cnt = rest;
                } else {
                    flush_buffer(st.td, st.ti, st.cnt, s_bc, lane, px, py, pz, cx, cy, cz, cid);
                    st.cnt = 0;
                }
                dbg_cycles(dbg, 11, t_m, lane);
                st.kth = kth_of(st.td, st.ti, k);
                s_qk[lane] = valid ? st.kth.d : -1.0;
            }
        }
        __syncthreads();
    }
    // merge what is left, so the caller's bounds are exact for what was seen
    if (__ballot(st.cnt > 0) != 0ull) {
        if (dbg && lane == 0) atomicAdd(&dbg[6], 1ull);  // end-of-list merges
        if (dbg && valid) atomicAdd(&dbg[13], (unsigned long long)st.cnt);
#ifdef PGALF_KNN_STATS
        st.acc += valid ? st.cnt : 0;
#endif
        const int wm = dbg ? wave_max_i(st.cnt) : 0;
        if (dbg && lane == 0) atomicAdd(&dbg[14], (unsigned long long)((wm + kBuf - 1) / kBuf));
        const unsigned long long t_m = dbg_clock(dbg);
        flush_buffer(st.td, st.ti, st.cnt, s_bc, lane, px, py, pz, cx, cy, cz, cid);
        dbg_cycles(dbg, 11, t_m, lane);
        st.cnt = 0;
        st.kth = kth_of(st.td, st.ti, k);
    }
    s_qk[lane] = valid ? st.kth.d : -1.0;
    __syncthreads();
    dbg_cycles(dbg, 10, t_scan, lane);
}

static_assert(sizeof(int) * kGroups * kFCap <= sizeof(Stage),
              "ring-sort scratch must fit in the staging area");

// Stable counting sort of a group's bucket list by ring (top bits), the
// group's kG lanes in parallel; tmp (>= kFCap ints of LDS) is scratch.  Group-uniform count.
__device__ inline void ring_sort(int *list, int count, int *tmp, int gshift, int gl) {
    const int per = (count + kG - 1) / kG;
    const int c0 = min(count, gl * per), c1 = min(count, c0 + per);
    int cnt[kRings];
#pragma unroll
    for (int q = 0; q < kRings; q++) cnt[q] = 0;
    for (int j = c0; j < c1; j++) {
        const int q = (unsigned)list[j] >> kRingShift;
#pragma unroll
        for (int r = 0; r < kRings; r++) cnt[r] += (q == r);
    }
    int base = 0;
#pragma unroll
    for (int q = 0; q < kRings; q++) {
        int incl = cnt[q];
#pragma unroll
        for (int o = 1; o < kG; o <<= 1) {
            int t = __shfl_up(incl, o, kG);
            if (gl >= o) incl += t;
        }
        int off = base + incl - cnt[q];
        base += __shfl(incl, gshift + kG - 1, 64);
        for (int j = c0; j < c1; j++)
            if (((unsigned)list[j] >> kRingShift) == (unsigned)q) tmp[off++] = list[j];
    }
    __syncthreads();
    for (int j = gl; j < count; j += kG) list[j] = tmp[j];
}

// One wavefront of the exact query: logical wave `wid` (queries
// [q0, q0 + 64) of a halo).
template <class P>
__device__ __forceinline__ void query_wave_exact(
    unsigned wid, const double *qx, const double *qy, const double *qz, const unsigned *qid,
    const double *cxa, const double *cya, const double *cza, const unsigned *cida,
    const Box *boxes, const KHalo *halos, const KHalo &halo0, const KWave *waves, int k,
    long long *nbr, float *rho, int rho_mode, const P *mass, int self_seed,
    unsigned long long *dbg, unsigned (*s_bc)[kWave], int (*s_list)[kFCap],
    int (*s_stk)[kStackCap], long long *s_off, long long *s_n, double *s_qx, double *s_qy,
    double *s_qz, double *s_qk, Stage &sg) {
    const unsigned long long t_start = dbg_clock(dbg);
    const int lane = threadIdx.x;
    const int grp = lane / kG, gl = lane % kG, gshift = grp * kG;
    // this wave's halo: its candidates, queries and tree
    KWave wv;
    if (waves) {
        wv = waves[wid];
    } else {
        wv.halo = 0;
        wv.q0 = (long long)wid * kWave;
    }
    const KHalo &hd = halos ? halos[wv.halo] : halo0;
    const long long nc = hd.nc, nb = hd.nb, hq0 = hd.q0, hq1 = hd.q0 + hd.nq;
    const long long idx_base = hd.idx_base;
    const int nlev = hd.nlev;
    const double *cx = cxa + hd.c0, *cy = cya + hd.c0, *cz = cza + hd.c0;
    const unsigned *cid = cida + hd.c0;
    if (lane < nlev) {
        s_off[lane] = hd.loff[lane];
        s_n[lane] = level_count(nb, lane);
    }
    const Box *leaves = boxes + hd.loff[0];
    const long long qi = wv.q0 + lane;
    const bool valid = qi < hq1;
    double px = 0, py = 0, pz = 0;
    if (valid) {
        px = qx[qi];
        py = qy[qi];
        pz = qz[qi];
    }
    double qlo[3], qhi[3];
    qlo[0] = grp_min_d(valid ? px : DBL_MAX);
    qlo[1] = grp_min_d(valid ? py : DBL_MAX);
    qlo[2] = grp_min_d(valid ? pz : DBL_MAX);
    qhi[0] = grp_max_d(valid ? px : -DBL_MAX);
    qhi[1] = grp_max_d(valid ? py : -DBL_MAX);
    qhi[2] = grp_max_d(valid ? pz : -DBL_MAX);
    const bool gvalid = ((__ballot(valid) >> gshift) & kGMask) != 0ull;

    QState st;
#pragma unroll
    for (int j = 0; j < kTop; j++) {
        st.td[j] = __longlong_as_double(0x7ff0000000000000ll);  // +inf
        st.ti[j] = 0xffffffffu;
    }
    st.kth.d = st.td[0];
    st.kth.i = st.ti[0];
    st.cnt = 0;
#ifdef PGALF_KNN_STATS
    st.acc = 0;
#endif
    __syncthreads();
    const double inf = __longlong_as_double(0x7ff0000000000000ll);

    // ---- seed bucket.  Queries = candidates: the group's own bucket (same
    // sorted order).  Otherwise descend greedily to the child whose box is
    // nearest the group's centre point (ties: nearest box centre).
    long long node = 0;
    if (self_seed) {
        node = (wv.q0 - hq0 + grp * kG) / kB;
    } else {
        const double gc[3] = {0.5 * (qlo[0] + qhi[0]), 0.5 * (qlo[1] + qhi[1]),
                              0.5 * (qlo[2] + qhi[2])};
        for (int l = nlev - 1; l > 0; l--) {
            long long child = kFan * node + (gl & (kFan - 1));
            bool ok = gl < kFan && child < s_n[l - 1];
            double g2 = inf, c2 = inf;
            if (ok) {
                Box bx = boxes[s_off[l - 1] + child];
                g2 = box_gap2(bx, gc, gc);
                double e0 = 0.5 * (bx.lo[0] + bx.hi[0]) - gc[0];
                double e1 = 0.5 * (bx.lo[1] + bx.hi[1]) - gc[1];
                double e2 = 0.5 * (bx.lo[2] + bx.hi[2]) - gc[2];
                c2 = e0 * e0 + e1 * e1 + e2 * e2;
            }
            long long cbest = ok ? child : 0x7fffffffffffffffll;
#pragma unroll
            for (int o = kG / 2; o > 0; o >>= 1) {
                double og = __shfl_xor(g2, o, 64);
                double oc2 = __shfl_xor(c2, o, 64);
                long long oc = __shfl_xor(cbest, o, 64);
                bool take = (og < g2) | ((og == g2) & ((oc2 < c2) | ((oc2 == c2) & (oc < cbest))));
                g2 = take ? og : g2;
                c2 = take ? oc2 : c2;
                cbest = take ? oc : cbest;
            }
            node = cbest;
        }
    }
    if (node > nb - 1) node = nb - 1;
    // direct fill: kSeedFill consecutive buckets from the one before the seed
    // bucket give 32 candidates, sorted once; the other seeds -- the window
    // [slo, shi] of kSeedSpan buckets around them -- go through the buffer.
    long long f0 = node - 1;
    if (f0 > nb - kSeedFill) f0 = nb - kSeedFill;
    if (f0 < 0) f0 = 0;
#pragma unroll
    for (int c = 0; c < kTop; c++) {
        const long long e = f0 * kB + c;
        if (gvalid && e < nc) {
            st.td[c] = d2_ref(px, py, pz, cx[e], cy[e], cz[e]);
            st.ti[c] = cid[e];
        }
    }
    oe_sort<kTop>(st.td, st.ti);
    st.kth = kth_of(st.td, st.ti, k);
    constexpr int kSeedBefore = (kSeedSpan - kSeedFill) / 2;
    const long long slo = f0 - kSeedBefore < 0 ? 0 : f0 - kSeedBefore;
    const long long shi = slo + kSeedSpan - 1 > nb - 1 ? nb - 1 : slo + kSeedSpan - 1;
    // the window's buckets next to the fill, [d0, f0) and [f0 + kSeedFill, d1),
    // merged straight into the sorted lists (no bound test, no buffer)
    const long long d0 = f0 - kDirect / 2 > slo ? f0 - kDirect / 2 : slo;
    const long long d1 = f0 + kSeedFill + kDirect / 2 < shi + 1 ? f0 + kSeedFill + kDirect / 2
                                                               : shi + 1;
    if constexpr (kDirect > 0) {
        const int nlo = (int)(f0 - d0);
        const int ndir = nlo + (d1 > f0 + kSeedFill ? (int)(d1 - f0 - kSeedFill) : 0);
#pragma unroll 1
        for (int c = 0; c < kDirect / 2; c++) {
            double bd[kDM];
            unsigned bi[kDM];
#pragma unroll
            for (int t = 0; t < kDM; t++) {
                const int j = 2 * c + t / kB;  // j-th direct bucket
                const long long b = j < nlo ? d0 + j : f0 + kSeedFill + (j - nlo);
                const long long e = b * kB + t % kB;
                bd[t] = inf;
                bi[t] = 0xffffffffu;
                if (gvalid && j < ndir && e < nc) {
                    bd[t] = d2_ref(px, py, pz, cx[e], cy[e], cz[e]);
                    bi[t] = cid[e];
                }
            }
            merge_top(st.td, st.ti, bd, bi);
        }
        st.kth = kth_of(st.td, st.ti, k);
    }
    // remaining seed buckets through the batched scanner (list in the stack)
    static_assert(kSeedSpan <= kStackCap, "seed list lives in the walk stack");
    int scnt = 0;
    if (gl == 0) {
        for (long long b = slo; b <= shi; b++)
            if (gvalid && (b < d0 || b >= d1)) s_stk[grp][scnt++] = (int)b;
    }
    scnt = __shfl(scnt, gshift, 64);
    __syncthreads();
    s_qx[lane] = px;
    s_qy[lane] = py;
    s_qz[lane] = pz;
    s_qk[lane] = valid ? st.kth.d : -1.0;
    __syncthreads();
    scan_list(st, s_stk[grp], scnt, valid, k, px, py, pz, cx, cy, cz, cid, nc, leaves, sg, s_bc,
              s_qx, s_qy, s_qz, s_qk, lane, grp, gl, dbg);
    double R2 = grp_max_d(valid ? st.kth.d : 0.0);
    dbg_cycles(dbg, 8, t_start, lane);
#ifdef PGALF_KNN_STATS
    const int acc_seed = st.acc;
#endif

    // ---- depth-first walk, nearest child first: a round pops one node per
    // group and its kG lanes test the node's children against the group box
    // and the union of the lanes' current balls.  Passing internal children
    // are pushed farthest first; passing buckets are appended to the group's
    // list, which is ring-sorted and scanned whenever some list nears
    // capacity, tightening the bounds for the rest of the walk.
    const double gcen[3] = {0.5 * (qlo[0] + qhi[0]), 0.5 * (qlo[1] + qhi[1]),
                            0.5 * (qlo[2] + qhi[2])};
    int *list = s_list[grp];
    int *stk = s_stk[grp];
    int sp = 0, lcnt = 0, nscan = 0;
    if (gvalid && nlev > 1) {
        if (gl == 0) stk[0] = nlev - 1;  // root: node 0 of the top level
        sp = 1;
    }
    __syncthreads();
    for (;;) {
        const bool more = sp > 0;
        const bool any = __ballot(more) != 0ull;
        if (any) {
            const unsigned long long t_r = dbg_clock(dbg);
            int ent = 0;
            if (more) ent = stk[sp - 1];
            sp -= more ? 1 : 0;
            const int cl = more ? (ent & ((1 << kLevelShift) - 1)) - 1 : 0;  // children's level
            const long long child = (long long)kFan * (ent >> kLevelShift) + gl;
            bool pass = false;
            double cg = inf;
            if (more && child < s_n[cl] && !(cl == 0 && child >= slo && child <= shi)) {
                Box cb = boxes[s_off[cl] + child];
                pass = !(box_gap2(cb, qlo, qhi) > R2) &&
                       union_pass(cb, grp, s_qx, s_qy, s_qz, s_qk);
                cg = box_gap2(cb, gcen, gcen);
            }
            const unsigned m = (unsigned)((__ballot(pass) >> gshift) & kGMask);
            const int npass = __popc(m);
            if (dbg && lane == 0) atomicAdd(&dbg[2], 1ull);  // walk rounds
            if (cl == 0) {
                int pos = lcnt + __popc(m & ((1u << gl) - 1u));
                // distance class: gap from the group centre in R/kRingSteps
                // steps, floor(kRingSteps sqrt(cg / R2)) capped at kRings - 1,
                // by comparisons
                const double cs = (double)(kRingSteps * kRingSteps) * cg;
                int ring = 0;
#pragma unroll
                for (int j = 1; j < kRings; j++) ring += cs >= (double)(j * j) * R2 ? 1 : 0;
                if (pass) list[pos] = (int)child | (ring << kRingShift);
                lcnt += npass;
            } else {
                // rank among the passing children by (gap, lane): nearer ones on top
                int nearer = 0;
#pragma unroll
                for (int j = 0; j < kG; j++) {
                    const double og = __shfl(cg, gshift + j, 64);
                    nearer += (((m >> j) & 1u) != 0u) & ((og < cg) | ((og == cg) & (j < gl)));
                }
                if (pass) stk[sp + npass - 1 - nearer] = (int)(child << kLevelShift) | cl;
                sp += npass;
            }
            __syncthreads();
            dbg_cycles(dbg, 9, t_r, lane);
        }
        const int lmax = wave_max_i(lcnt);
        if (lmax > kFCap - kG || (!any && lmax > 0) || (nscan == 0 && lmax >= kFirstScan)) {
            nscan++;
            ring_sort(list, lcnt, reinterpret_cast<int *>(&sg) + grp * kFCap, gshift, gl);
            __syncthreads();
            if (dbg && lane == 0) atomicAdd(&dbg[4], 1ull);  // list scans
            int skip = 0;
            if constexpr (kWalkDirect > 0) {
                if (nscan == 1) {
                    // the first list's nearest buckets merged straight in
                    skip = lcnt < kWalkDirect ? lcnt : kWalkDirect;
#pragma unroll 1
                    for (int c = 0; c < kWalkDirect / 2; c++) {
                        double bd[kDM];
                        unsigned bi[kDM];
#pragma unroll
                        for (int t = 0; t < kDM; t++) {
                            const int j = 2 * c + t / kB;
                            bd[t] = inf;
                            bi[t] = 0xffffffffu;
                            if (j < skip) {
                                const long long e = (long long)(list[j] & kBucketMask) * kB + t % kB;
                                if (e < nc) {
                                    bd[t] = d2_ref(px, py, pz, cx[e], cy[e], cz[e]);
                                    bi[t] = cid[e];
                                }
                            }
                        }
                        merge_top(st.td, st.ti, bd, bi);
                    }
                    st.kth = kth_of(st.td, st.ti, k);
                    s_qk[lane] = valid ? st.kth.d : -1.0;
                    __syncthreads();
                }
            }
            scan_list(st, list + skip, lcnt - skip, valid, k, px, py, pz, cx, cy, cz, cid, nc,
                      leaves, sg, s_bc, s_qx, s_qy, s_qz, s_qk, lane, grp, gl, dbg);
            lcnt = 0;
            R2 = grp_max_d(valid ? st.kth.d : 0.0);
        }
        if (!any) break;
    }
    if (dbg && lane == 0) atomicAdd(&dbg[5], 1ull);  // waves
#ifdef PGALF_KNN_STATS
    if (dbg) {
        const int mx = wave_max_i(st.acc), mxs = wave_max_i(acc_seed);
        const int gmx = grp_max_i(st.acc);
        if (valid) {
            atomicAdd(&dbg[16], (unsigned long long)acc_seed);
            atomicAdd(&dbg[17], (unsigned long long)(st.acc - acc_seed));
            atomicAdd(&dbg[24 + min(st.acc / 16, 31)], 1ull);
        }
        if (lane == 0) {
            atomicAdd(&dbg[18], (unsigned long long)mx);
            atomicAdd(&dbg[19], (unsigned long long)mxs);
        }
        if (gl == 0) atomicAdd(&dbg[20], (unsigned long long)gmx);
    }
#endif
    dbg_cycles(dbg, 12, t_start, lane);
    if (!valid) return;
    const long long q = qid[qi];
    if (nbr) {
        long long *row = nbr + q * k;
#pragma unroll
        for (int t = 0; t < kTop; t++)
            if (t < k) row[t] = (long long)st.ti[t] - idx_base;
    }
    if (rho) {
        float neardist = sqrtf((float)st.kth.d);
        float h = (float)((double)neardist / 2.0);
        float sum = 0.0f;
#pragma unroll
        for (int t = 0; t < kTop; t++) {
            if (t < k) {
                float r = sqrtf((float)st.td[t]);
                float w = W4(r, h);
                if (rho_mode == 2) w = w * (float)mass[st.ti[t]];
                sum += w;
            }
        }
        rho[q] = sum;
    }
}

// The exact query kernel.  kList == false: block b runs logical wave
// xcd_block(b, gridDim.x).  kList == true: the blocks run the waves listed in
// wlist[0 .. *wcount) (those the packed-key kernel flagged), grid-stride --
// a separate instantiation, so the loop does not cost the plain kernel
// registers.  The list instantiation runs a few blocks at most and asks for
// one wave per SIMD, so what does not fit 256 VGPRs goes to AGPRs rather
// than to scratch memory (a dispatch that needs scratch costs far more than
// this kernel's own work when its list is empty).
template <class P, bool kList>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(kList ? 1 : 2)))
knn_query_k(const double *qx, const double *qy, const double *qz, const unsigned *qid,
            const double *cxa, const double *cya, const double *cza, const unsigned *cida,
            const Box *boxes, const KHalo *halos, KHalo halo0, const KWave *waves, int k,
            long long *nbr, float *rho, int rho_mode, const P *mass,
            int self_seed, unsigned long long *dbg, const unsigned *wlist,
            const unsigned *wcount, unsigned xcd_chunk) {
    __shared__ unsigned s_bc[kBufCap][kWave];
    __shared__ int s_list[kGroups][kFCap];
    __shared__ int s_stk[kGroups][kStackCap];  // also the seed list before the walk
    __shared__ long long s_off[kMaxLevels], s_n[kMaxLevels];
    __shared__ double s_qx[kWave], s_qy[kWave], s_qz[kWave], s_qk[kWave];
    __shared__ Stage sg;
    if constexpr (kList) {
        const unsigned n = *wcount;
        for (unsigned it = blockIdx.x; it < n; it += gridDim.x) {
            query_wave_exact<P>(wlist[it], qx, qy, qz, qid, cxa, cya, cza, cida, boxes, halos,
                                halo0, waves, k, nbr, rho, rho_mode, mass, self_seed, dbg, s_bc,
                                s_list, s_stk, s_off, s_n, s_qx, s_qy, s_qz, s_qk, sg);
            __syncthreads();
        }
    } else {
        query_wave_exact<P>(xcd_block(blockIdx.x, gridDim.x, xcd_chunk), qx, qy, qz, qid, cxa, cya, cza,
                            cida, boxes, halos, halo0, waves, k, nbr, rho, rho_mode, mass,
                            self_seed, dbg, s_bc, s_list, s_stk, s_off, s_n, s_qx, s_qy, s_qz,
                            s_qk, sg);
    }
}

// ---- packed-key query --------------------------------------------------------
// The same search as the exact query (seed window, depth-first walk, batched
// list scans, per-lane buffers merged into a sorted top-32 in registers),
// with each candidate's (distance, id) pair packed into one double:
//
//   key bits = ((bits(d2) >> 1) + 2^52) with the low kb bits replaced by the
//              candidate's slot e (nb * kB <= 2^kb).
//
// For d2 >= 0 the bits of d2 are monotone in d2, so the key orders
// candidates by d2 truncated to 51 - kb bits (its "class") and, within a
// class, by slot; every key is a positive normal double.  A compare-exchange
// is then a v_min_f64 / v_max_f64 pair (no lexicographic select), and the
// list is 32 doubles.
//
// Exactness.  Outside one class, key order is the exact (d2, id) order.  A
// query's key-ordered top k can differ from the exact top k only if a
// displaced candidate shares the k-th key's class, or two of the first k
// entries share a class.  The scan accepts every candidate whose d2 is at
// most the largest d2 of the current k-th key's class (class_ub), so every
// candidate of the final k-th class that is ever seen passes through a merge;
// merges record the smallest key they displace (L).  Candidates of a pruned
// box or rejected by the scan lie in a larger class than the k-th.  A wave
// with a lane that fails the check (same_class(L, k-th) or equal classes
// among the first k) writes nothing and is listed in fix_list; the exact
// query then recomputes those waves.  For distinct float positions such
// lanes are about one in 10^6; duplicated points and lattices flag many.
constexpr unsigned long long kKeyOff = 1ull << 52;
// A wave with at most kRadLanes flagged lanes writes its other lanes and
// lists the flagged ones for knn_radius_k; a wave with more is listed whole.
constexpr int kRadLanes = 4;
__device__ inline double key_inf() { return __longlong_as_double(0x7fe0000000000000ll); }

__device__ inline double pkey(double d2, unsigned e, unsigned long long keep) {
    unsigned long long b = (unsigned long long)__double_as_longlong(d2);
    b = (((b >> 1) + kKeyOff) & keep) | e;
    return __longlong_as_double((long long)b);
}
__device__ inline unsigned pslot(double key, unsigned long long keep) {
    return (unsigned)((unsigned long long)__double_as_longlong(key) & ~keep);
}
__device__ inline bool same_class(double a, double b, unsigned long long keep) {
    return ((((unsigned long long)__double_as_longlong(a)) ^
             ((unsigned long long)__double_as_longlong(b))) & keep) == 0ull;
}
// The largest d2 whose key has key's class: every d2 of a larger class
// exceeds it (+inf for the empty-slot key).
__device__ inline double class_ub(double key, unsigned long long keep) {
    if (!(key < key_inf())) return __longlong_as_double(0x7ff0000000000000ll);
    unsigned long long b = (unsigned long long)__double_as_longlong(key);
    b = (((b - kKeyOff) | ~keep) << 1) | 1ull;
    return __longlong_as_double((long long)b);
}

__device__ inline void pcas(double &a, double &b) {
    const double lo = fmin(a, b), hi = fmax(a, b);
    a = lo;
    b = hi;
}
template <int N, int LO, int HI, int R>
__device__ inline void oe_merge_p(double (&d)[N]) {
    constexpr int STEP = 2 * R;
    if constexpr (STEP < HI - LO) {
        oe_merge_p<N, LO, HI, STEP>(d);
        oe_merge_p<N, LO + R, HI, STEP>(d);
#pragma unroll
        for (int i = LO + R; i < HI - R; i += STEP) pcas(d[i], d[i + R]);
    } else {
        pcas(d[LO], d[LO + R]);
    }
}
template <int N, int LO = 0, int HI = N - 1>
__device__ inline void oe_sort_p(double (&d)[N]) {
    if constexpr (HI - LO >= 1) {
        constexpr int MID = LO + (HI - LO) / 2;
        oe_sort_p<N, LO, MID>(d);
        oe_sort_p<N, MID + 1, HI>(d);
        oe_merge_p<N, LO, HI, 1>(d);
    }
}
template <int N, int J>
__device__ inline void bitonic_clean_p(double (&d)[N]) {
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int l = i ^ J;
        if (l > i) pcas(d[i], d[l]);
    }
    if constexpr (J > 1) bitonic_clean_p<N, J / 2>(d);
}
// Merge a chunk of N keys into the sorted top-32 (as merge_top); the keys
// that leave the list lower L.
template <int N>
__device__ __forceinline__ void merge_top_p(double (&td)[kTop], double (&bd)[N], double &L) {
    oe_sort_p<N>(bd);
#pragma unroll
    for (int j = 0; j < N; j++) {
        constexpr int o = kTop - N;
        const double c = bd[N - 1 - j], t = td[o + j];
        td[o + j] = fmin(t, c);
        L = fmin(L, fmax(t, c));
    }
    bitonic_clean_p<kTop, kTop / 2>(td);
}

// Per-phase shader-clock sums of the packed-key query (PGALF_KNN_PROF=1; a
// separate instantiation, so the plain kernel carries none of it).
enum : int {
    kPrSeed = 0, kPrWalk, kPrScan, kPrMerge, kPrTail, kPrTotal,
    kPrWaves, kPrRounds, kPrBatches, kPrChunks, kPrFlagged, kPrSlots, kPrN = 16
};
template <bool kProf>
__device__ inline unsigned long long pr_clock() {
    if constexpr (kProf) return (unsigned long long)clock64();
    return 0ull;
}
template <bool kProf>
__device__ inline void pr_add(unsigned long long *prof, int i, unsigned long long v, int lane) {
    if constexpr (kProf) {
        if (lane == 0) atomicAdd(&prof[i], v);
    }
}

struct PState {
    double td[kTop];  // sorted keys
    double kth;       // td[k - 1]
    double kub;       // class_ub(kth); -1 for an invalid lane (accepts nothing)
    double L;         // smallest key displaced from the list
    int cnt;
};

__device__ __forceinline__ void p_bound(PState &st, int k, bool valid, unsigned long long keep) {
    double r = st.td[kTop - 1];
    if (k != kTop) {
        // sorted: the (k-1)-th is the largest of the first k, taken as a
        // masked maximum -- a masked select would be folded into a dynamic
        // index, which puts the list in scratch memory
        r = st.td[0];
#pragma unroll
        for (int t = 1; t < kTop; t++) r = t < k ? fmax(r, st.td[t]) : r;
    }
    st.kth = r;
    st.kub = valid ? class_ub(r, keep) : -1.0;
}

// Gaps as max(lo - q, q - hi, 0): the same value as gap1 without branches.
__device__ inline double gap1_p(double lo, double hi, double qlo, double qhi) {
    return fmax(fmax(lo - qhi, qlo - hi), 0.0);
}
__device__ __forceinline__ double box_gap2_p(const Box &b, const double qlo[3], const double qhi[3]) {
    const double g0 = gap1_p(b.lo[0], b.hi[0], qlo[0], qhi[0]);
    const double g1 = gap1_p(b.lo[1], b.hi[1], qlo[1], qhi[1]);
    const double g2 = gap1_p(b.lo[2], b.hi[2], qlo[2], qhi[2]);
    const double a = g0 * g0;
    const double bb = g1 * g1;
    const double c = g2 * g2;
    return (a + bb) + c;
}
// union_pass with the test folded into a minimum: g2 <= qk exactly when
// g2 - qk <= 0 (IEEE subtraction keeps the sign and is zero only for equal
// operands); invalid lanes hold qk = -1 and never pass.
__device__ __forceinline__ bool union_pass_p(const Box &b, int grp, const double *s_qx,
                                    const double *s_qy, const double *s_qz, const double *s_qk) {
    double mn = __longlong_as_double(0x7ff0000000000000ll);
#pragma unroll
    for (int j = 0; j < kG; j++) {
        const int L = grp * kG + j;
        const double x = s_qx[L], y = s_qy[L], z = s_qz[L];
        const double g0 = gap1_p(b.lo[0], b.hi[0], x, x);
        const double g1 = gap1_p(b.lo[1], b.hi[1], y, y);
        const double g2 = gap1_p(b.lo[2], b.hi[2], z, z);
        const double a = g0 * g0;
        const double bb = g1 * g1;
        const double c = g2 * g2;
        mn = fmin(mn, ((a + bb) + c) - s_qk[L]);
    }
    return mn <= 0.0;
}

// Merge a lane's buffer of cnt sorted-candidate slots into its top-32, nch
// chunks of kBuf at most (wave uniform); keys recomputed from the slots (a
// buffer of keys would be twice the LDS, or half the slots and twice the
// flushes: both measured slower).
template <bool kProf>
__device__ __forceinline__ void flush_p(PState &st, int nch, const unsigned (*s_bc)[kWave], int lane,
                               double px, double py, double pz, const double *cx,
                               const double *cy, const double *cz, unsigned long long keep,
                               unsigned long long *prof) {
    const unsigned long long t0 = pr_clock<kProf>();
#pragma unroll 1
    for (int ch = 0; ch < nch; ch++) {
        if (__ballot(st.cnt > ch * kBuf) == 0ull) break;
        pr_add<kProf>(prof, kPrChunks, 1ull, lane);
        double bd[kBuf];
#pragma unroll
        for (int t = 0; t < kBuf; t++) {
            const int e = ch * kBuf + t;
            bd[t] = key_inf();
            if (e < st.cnt) {
                const unsigned sidx = s_bc[e][lane];
                bd[t] = pkey(d2_ref(px, py, pz, cx[sidx], cy[sidx], cz[sidx]), sidx, keep);
            }
        }
        merge_top_p(st.td, bd, st.L);
    }
    pr_add<kProf>(prof, kPrMerge, pr_clock<kProf>() - t0, lane);
}

// scan_list of the packed-key query: a candidate goes to the buffer when its
// d2 is at most the lane's kub.
template <bool kProf>
__device__ __forceinline__ void scan_list_p(PState &st, int *list, int count, bool valid, int k,
                                   double px, double py, double pz, const double *cx,
                                   const double *cy, const double *cz, long long nc,
                                   const Box *leaves, Stage &sg, unsigned (*s_bc)[kWave],
                                   double *s_qx, double *s_qy, double *s_qz, double *s_qk,
                                   int lane, int grp, int gl, unsigned long long keep,
                                   unsigned long long *prof) {
    const double nan = __longlong_as_double(0x7ff8000000000000ll);
    int ipos = 0;
    const int nit = wave_max_i((count + kBatch - 1) / kBatch);
    for (int it = 0; it < nit; it++) {
        if (__ballot(ipos < count) == 0ull) break;
        const int wend = ipos + kG < count ? ipos + kG : count;
        int ent = 0;
        bool pass = false;
        if (ipos + gl < count) {
            ent = list[ipos + gl];
            pass = union_pass_p(leaves[ent & kBucketMask], grp, s_qx, s_qy, s_qz, s_qk);
        }
        const unsigned mm = (unsigned)((__ballot(pass) >> (grp * kG)) & kGMask);
        const int npass = __popc(mm);
        const int first = wend - npass;
        if (pass) list[first + __popc(mm & ((1u << gl) - 1u))] = ent;
        const int nsel = npass < kBatch ? npass : kBatch;
        if (ipos < count) ipos = first + nsel;
        double vx[kBatch], vy[kBatch], vz[kBatch];
        unsigned ve[kBatch];
#pragma unroll
        for (int t = 0; t < kBatch; t++) {
            long long e = t < nsel ? (long long)(list[first + t] & kBucketMask) * kB + gl : -1;
            if (e >= nc) e = -1;
            vx[t] = nan;
            vy[t] = 0.0;
            vz[t] = 0.0;
            ve[t] = 0xffffffffu;
            if (e >= 0) {
                vx[t] = cx[e];
                vy[t] = cy[e];
                vz[t] = cz[e];
                ve[t] = (unsigned)e;
            }
        }
        const int wsel = wave_max_i(nsel);
        pr_add<kProf>(prof, kPrBatches, 1ull, lane);
        pr_add<kProf>(prof, kPrSlots, (unsigned long long)wsel, lane);
#pragma unroll
        for (int t = 0; t < kBatch; t++)
            stage_put(sg, grp, t * kB + gl, vx[t], vy[t], vz[t], ve[t], ve[t]);
        __syncthreads();
        const int nslots = wsel * kB;
        const uint4 *su = sg.u + grp * kStageStride;
        for (int h = 0; h < nslots; h += kChunkC) {
#pragma unroll
            for (int c = h; c < h + kChunkC; c++) {
                const uint4 a = su[2 * c], b = su[2 * c + 1];
                const double d = d2_ref(px, py, pz, hilo_d(a.x, a.y), hilo_d(a.z, a.w),
                                        hilo_d(b.x, b.y));
                s_bc[st.cnt][lane] = b.z;
                st.cnt += d <= st.kub ? 1 : 0;
            }
            if (__ballot(st.cnt > kBufCap - kChunkC) != 0ull) {
                // one chunk for every lane; a lane's slots past it move down
                flush_p<kProf>(st, 1, s_bc, lane, px, py, pz, cx, cy, cz, keep, prof);
                const int rest = st.cnt > kBuf ? st.cnt - kBuf : 0;
                const int wrest = wave_max_i(rest);
                for (int t = 0; t < wrest; t++)
                    if (t < rest) s_bc[t][lane] = s_bc[kBuf + t][lane];
                st.cnt = rest;
                p_bound(st, k, valid, keep);
                s_qk[lane] = st.kub;
            }
        }
        __syncthreads();
    }
    if (__ballot(st.cnt > 0) != 0ull) {
        flush_p<kProf>(st, kBufCap / kBuf, s_bc, lane, px, py, pz, cx, cy, cz, keep, prof);
        st.cnt = 0;
        p_bound(st, k, valid, keep);
    }
    s_qk[lane] = st.kub;
    __syncthreads();
}

// One wavefront of the packed-key query: logical wave `wid`.
template <class P, bool kProf>
__device__ __forceinline__ void packed_wave(
    unsigned wid, const double *qx, const double *qy, const double *qz, const unsigned *qid,
    const double *cxa, const double *cya, const double *cza, const unsigned *cida,
    const Box *boxes, const KHalo *halos, const KHalo &halo0, const KWave *waves, int k,
    long long *nbr, float *rho, int rho_mode, const P *mass, int self_seed,
    unsigned *fix_count, unsigned *fix_list, unsigned *rad, double *rad_r2,
    unsigned long long *prof,
    unsigned (*s_bc)[kWave], int (*s_list)[kFCap], int (*s_stk)[kStackCap], long long *s_off,
    long long *s_n, double *s_qx, double *s_qy, double *s_qz, double *s_qk, Stage &sg) {
    const unsigned long long t_start = pr_clock<kProf>();
    const int lane = threadIdx.x;
    const int grp = lane / kG, gl = lane % kG, gshift = grp * kG;
    KWave wv;
    if (waves) {
        wv = waves[wid];
    } else {
        wv.halo = 0;
        wv.q0 = (long long)wid * kWave;
    }
    const KHalo &hd = halos ? halos[wv.halo] : halo0;
    const long long nc = hd.nc, nb = hd.nb, hq0 = hd.q0, hq1 = hd.q0 + hd.nq;
    const long long idx_base = hd.idx_base;
    const int nlev = hd.nlev;
    const unsigned long long keep = ~((1ull << hd.kb) - 1ull);
    const double *cx = cxa + hd.c0, *cy = cya + hd.c0, *cz = cza + hd.c0;
    const unsigned *cid = cida + hd.c0;
    if (lane < nlev) {
        s_off[lane] = hd.loff[lane];
        s_n[lane] = level_count(nb, lane);
    }
    const Box *leaves = boxes + hd.loff[0];
    const long long qi = wv.q0 + lane;
    const bool valid = qi < hq1;
    double px = 0, py = 0, pz = 0;
    if (valid) {
        px = qx[qi];
        py = qy[qi];
        pz = qz[qi];
    }
    double qlo[3], qhi[3];
    qlo[0] = grp_min_d(valid ? px : DBL_MAX);
    qlo[1] = grp_min_d(valid ? py : DBL_MAX);
    qlo[2] = grp_min_d(valid ? pz : DBL_MAX);
    qhi[0] = grp_max_d(valid ? px : -DBL_MAX);
    qhi[1] = grp_max_d(valid ? py : -DBL_MAX);
    qhi[2] = grp_max_d(valid ? pz : -DBL_MAX);
    const bool gvalid = ((__ballot(valid) >> gshift) & kGMask) != 0ull;

    PState st;
    const double kinf = key_inf();
#pragma unroll
    for (int j = 0; j < kTop; j++) st.td[j] = kinf;
    st.L = kinf;
    st.cnt = 0;
    __syncthreads();
    const double inf = __longlong_as_double(0x7ff0000000000000ll);

    // ---- seed bucket (as the exact query)
    long long node = 0;
    if (self_seed) {
        node = (wv.q0 - hq0 + grp * kG) / kB;
    } else {
        const double gc[3] = {0.5 * (qlo[0] + qhi[0]), 0.5 * (qlo[1] + qhi[1]),
                              0.5 * (qlo[2] + qhi[2])};
        for (int l = nlev - 1; l > 0; l--) {
            long long child = kFan * node + (gl & (kFan - 1));
            bool ok = gl < kFan && child < s_n[l - 1];
            double g2 = inf, c2 = inf;
            if (ok) {
                Box bx = boxes[s_off[l - 1] + child];
                g2 = box_gap2(bx, gc, gc);
                double e0 = 0.5 * (bx.lo[0] + bx.hi[0]) - gc[0];
                double e1 = 0.5 * (bx.lo[1] + bx.hi[1]) - gc[1];
                double e2 = 0.5 * (bx.lo[2] + bx.hi[2]) - gc[2];
                c2 = e0 * e0 + e1 * e1 + e2 * e2;
            }
            long long cbest = ok ? child : 0x7fffffffffffffffll;
#pragma unroll
            for (int o = kG / 2; o > 0; o >>= 1) {
                double og = __shfl_xor(g2, o, 64);
                double oc2 = __shfl_xor(c2, o, 64);
                long long oc = __shfl_xor(cbest, o, 64);
                bool take = (og < g2) | ((og == g2) & ((oc2 < c2) | ((oc2 == c2) & (oc < cbest))));
                g2 = take ? og : g2;
                c2 = take ? oc2 : c2;
                cbest = take ? oc : cbest;
            }
            node = cbest;
        }
    }
    if (node > nb - 1) node = nb - 1;
    long long f0 = node - 1;
    if (f0 > nb - kSeedFill) f0 = nb - kSeedFill;
    if (f0 < 0) f0 = 0;
#pragma unroll
    for (int c = 0; c < kTop; c++) {
        const long long e = f0 * kB + c;
        if (gvalid && e < nc)
            st.td[c] = pkey(d2_ref(px, py, pz, cx[e], cy[e], cz[e]), (unsigned)e, keep);
    }
    oe_sort_p<kTop>(st.td);
    constexpr int kSeedBefore = (kSeedSpan - kSeedFill) / 2;
    const long long slo = f0 - kSeedBefore < 0 ? 0 : f0 - kSeedBefore;
    const long long shi = slo + kSeedSpan - 1 > nb - 1 ? nb - 1 : slo + kSeedSpan - 1;
    const long long d0 = f0 - kDirect / 2 > slo ? f0 - kDirect / 2 : slo;
    const long long d1 = f0 + kSeedFill + kDirect / 2 < shi + 1 ? f0 + kSeedFill + kDirect / 2
                                                               : shi + 1;
    if constexpr (kDirect > 0) {
        const unsigned long long t_m = pr_clock<kProf>();
        const int nlo = (int)(f0 - d0);
        const int ndir = nlo + (d1 > f0 + kSeedFill ? (int)(d1 - f0 - kSeedFill) : 0);
#pragma unroll 1
        for (int c = 0; c < kDirect / 2; c++) {
            double bd[kDM];
#pragma unroll
            for (int t = 0; t < kDM; t++) {
                const int j = 2 * c + t / kB;
                const long long b = j < nlo ? d0 + j : f0 + kSeedFill + (j - nlo);
                const long long e = b * kB + t % kB;
                bd[t] = kinf;
                if (gvalid && j < ndir && e < nc)
                    bd[t] = pkey(d2_ref(px, py, pz, cx[e], cy[e], cz[e]), (unsigned)e, keep);
            }
            merge_top_p(st.td, bd, st.L);
        }
        pr_add<kProf>(prof, kPrMerge, pr_clock<kProf>() - t_m, lane);
    }
    p_bound(st, k, valid, keep);
    int scnt = 0;
    if (gl == 0) {
        for (long long b = slo; b <= shi; b++)
            if (gvalid && (b < d0 || b >= d1)) s_stk[grp][scnt++] = (int)b;
    }
    scnt = __shfl(scnt, gshift, 64);
    __syncthreads();
    s_qx[lane] = px;
    s_qy[lane] = py;
    s_qz[lane] = pz;
    s_qk[lane] = st.kub;
    __syncthreads();
    scan_list_p<kProf>(st, s_stk[grp], scnt, valid, k, px, py, pz, cx, cy, cz, nc, leaves, sg,
                       s_bc, s_qx, s_qy, s_qz, s_qk, lane, grp, gl, keep, prof);
    double R2 = grp_max_d(valid ? st.kub : 0.0);
    pr_add<kProf>(prof, kPrSeed, pr_clock<kProf>() - t_start, lane);

    // ---- depth-first walk, nearest child first (as the exact query)
    const double gcen[3] = {0.5 * (qlo[0] + qhi[0]), 0.5 * (qlo[1] + qhi[1]),
                            0.5 * (qlo[2] + qhi[2])};
    int *list = s_list[grp];
    int *stk = s_stk[grp];
    int sp = 0, lcnt = 0, nscan = 0;
    if (gvalid && nlev > 1) {
        if (gl == 0) stk[0] = nlev - 1;
        sp = 1;
    }
    __syncthreads();
    for (;;) {
        const bool more = sp > 0;
        const bool any = __ballot(more) != 0ull;
        if (any) {
            const unsigned long long t_r = pr_clock<kProf>();
            pr_add<kProf>(prof, kPrRounds, 1ull, lane);
            int ent = 0;
            if (more) ent = stk[sp - 1];
            sp -= more ? 1 : 0;
            const int cl = more ? (ent & ((1 << kLevelShift) - 1)) - 1 : 0;
            const long long child = (long long)kFan * (ent >> kLevelShift) + gl;
            bool pass = false;
            double cg = inf;
            if (more && child < s_n[cl] && !(cl == 0 && child >= slo && child <= shi)) {
                Box cb = boxes[s_off[cl] + child];
                pass = !(box_gap2_p(cb, qlo, qhi) > R2) &&
                       union_pass_p(cb, grp, s_qx, s_qy, s_qz, s_qk);
                cg = box_gap2_p(cb, gcen, gcen);
            }
            const unsigned m = (unsigned)((__ballot(pass) >> gshift) & kGMask);
            const int npass = __popc(m);
            if (cl == 0) {
                int pos = lcnt + __popc(m & ((1u << gl) - 1u));
                const double cs = (double)(kRingSteps * kRingSteps) * cg;
                int ring = 0;
#pragma unroll
                for (int j = 1; j < kRings; j++) ring += cs >= (double)(j * j) * R2 ? 1 : 0;
                if (pass) list[pos] = (int)child | (ring << kRingShift);
                lcnt += npass;
            } else {
                int nearer = 0;
#pragma unroll
                for (int j = 0; j < kG; j++) {
                    const double og = __shfl(cg, gshift + j, 64);
                    nearer += (((m >> j) & 1u) != 0u) & ((og < cg) | ((og == cg) & (j < gl)));
                }
                if (pass) stk[sp + npass - 1 - nearer] = (int)(child << kLevelShift) | cl;
                sp += npass;
            }
            __syncthreads();
            pr_add<kProf>(prof, kPrWalk, pr_clock<kProf>() - t_r, lane);
        }
        const int lmax = wave_max_i(lcnt);
        if (lmax > kFCap - kG || (!any && lmax > 0) || (nscan == 0 && lmax >= kFirstScan)) {
            const unsigned long long t_s = pr_clock<kProf>();
            nscan++;
            ring_sort(list, lcnt, reinterpret_cast<int *>(&sg) + grp * kFCap, gshift, gl);
            __syncthreads();
            int skip = 0;
            if constexpr (kWalkDirect > 0) {
                if (nscan == 1) {
                    const unsigned long long t_m = pr_clock<kProf>();
                    skip = lcnt < kWalkDirect ? lcnt : kWalkDirect;
#pragma unroll 1
                    for (int c = 0; c < kWalkDirect / 2; c++) {
                        double bd[kDM];
#pragma unroll
                        for (int t = 0; t < kDM; t++) {
                            const int j = 2 * c + t / kB;
                            bd[t] = kinf;
                            if (j < skip) {
                                const long long e = (long long)(list[j] & kBucketMask) * kB + t % kB;
                                if (e < nc)
                                    bd[t] = pkey(d2_ref(px, py, pz, cx[e], cy[e], cz[e]),
                                                 (unsigned)e, keep);
                            }
                        }
                        merge_top_p(st.td, bd, st.L);
                    }
                    p_bound(st, k, valid, keep);
                    s_qk[lane] = st.kub;
                    __syncthreads();
                    pr_add<kProf>(prof, kPrMerge, pr_clock<kProf>() - t_m, lane);
                }
            }
            scan_list_p<kProf>(st, list + skip, lcnt - skip, valid, k, px, py, pz, cx, cy, cz,
                               nc, leaves, sg, s_bc, s_qx, s_qy, s_qz, s_qk, lane, grp, gl,
                               keep, prof);
            lcnt = 0;
            R2 = grp_max_d(valid ? st.kub : 0.0);
            pr_add<kProf>(prof, kPrScan, pr_clock<kProf>() - t_s, lane);
        }
        if (!any) break;
    }
    const unsigned long long t_tail = pr_clock<kProf>();

    // ---- exactness check (see the block comment): list entries past k are
    // displaced keys too
    double Lf = st.L;
#pragma unroll
    for (int t = 1; t < kTop; t++) Lf = t >= k ? fmin(Lf, st.td[t]) : Lf;
    bool flag = same_class(Lf, st.kth, keep);
#pragma unroll
    for (int t = 0; t + 1 < kTop; t++)
        flag |= (t + 1 < k) & same_class(st.td[t], st.td[t + 1], keep);
    flag &= valid;
    pr_add<kProf>(prof, kPrWaves, 1ull, lane);
    const unsigned long long fm = __ballot(flag);
    if (fm != 0ull) {
        pr_add<kProf>(prof, kPrFlagged, 1ull, lane);
        if (__popcll(fm) > kRadLanes) {
            // many lanes (duplicated points, lattices): the whole wave again
            if (lane == 0) fix_list[atomicAdd(fix_count, 1u)] = wid;
            return;
        }
        if (flag) {
            // this lane alone, by the radius of its k-th key's class
            const unsigned i = atomicAdd(rad, 1u);
            rad[1 + i] = wid * kWave + lane;
            rad_r2[i] = class_ub(st.kth, keep);
            return;
        }
    }
    if constexpr (kProf) {
        // the tail and total of unflagged waves, measured at their end
        const unsigned long long t_end = pr_clock<kProf>();
        pr_add<kProf>(prof, kPrTail, t_end - t_tail, lane);
        pr_add<kProf>(prof, kPrTotal, t_end - t_start, lane);
    }
    if (!valid) return;
    const long long q = qid[qi];
    if (nbr) {
        long long *row = nbr + q * k;
#pragma unroll
        for (int t = 0; t < kTop; t++)
            if (t < k) row[t] = (long long)cid[pslot(st.td[t], keep)] - idx_base;
    }
    if (rho) {
        // exact FP64 distances again from the slots, summed as the exact query does
        double dk = 0.0;
        {
            const unsigned e = pslot(st.kth, keep);
            dk = d2_ref(px, py, pz, cx[e], cy[e], cz[e]);
        }
        const float neardist = sqrtf((float)dk);
        const float h = (float)((double)neardist / 2.0);
        float sum = 0.0f;
#pragma unroll
        for (int t = 0; t < kTop; t++) {
            if (t < k) {
                const unsigned e = pslot(st.td[t], keep);
                const float r = sqrtf((float)d2_ref(px, py, pz, cx[e], cy[e], cz[e]));
                float w = W4(r, h);
                if (rho_mode == 2) w = w * (float)mass[cid[e]];
                sum += w;
            }
        }
        rho[q] = sum;
    }
}

#define PGALF_PACKED_ARGS                                                                    \
    qx, qy, qz, qid, cxa, cya, cza, cida, boxes, halos, halo0, waves, k, nbr, rho, rho_mode,  \
        mass, self_seed, fix_count, fix_list, rad, rad_r2, prof, s_bc, s_list, s_stk, s_off,  \
        s_n, s_qx,                                                                           \
        s_qy, s_qz, s_qk, sg
#define PGALF_PACKED_SHARED                                                                  \
    __shared__ unsigned s_bc[kBufCap][kWave];                                                \
    __shared__ int s_list[kGroups][kFCap];                                                   \
    __shared__ int s_stk[kGroups][kStackCap]; /* also the seed list before the walk */       \
    __shared__ long long s_off[kMaxLevels], s_n[kMaxLevels];                                 \
    __shared__ double s_qx[kWave], s_qy[kWave], s_qz[kWave], s_qk[kWave];                    \
    __shared__ Stage sg;

// The packed-key query kernel: block b runs logical wave xcd_block(b, gridDim.x).
template <class P, bool kProf>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2)))
knn_packed_k(const double *qx, const double *qy, const double *qz, const unsigned *qid,
             const double *cxa, const double *cya, const double *cza, const unsigned *cida,
             const Box *boxes, const KHalo *halos, KHalo halo0, const KWave *waves, int k,
             long long *nbr, float *rho, int rho_mode, const P *mass, int self_seed,
             unsigned *fix_count, unsigned *fix_list, unsigned *rad, double *rad_r2,
             unsigned long long *prof, unsigned xcd_chunk) {
    PGALF_PACKED_SHARED
    packed_wave<P, kProf>(xcd_block(blockIdx.x, gridDim.x, xcd_chunk), PGALF_PACKED_ARGS);
}

#undef PGALF_PACKED_ARGS
#undef PGALF_PACKED_SHARED

inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// Blocks of the re-runs, grid-stride over their lists (short for distinct
// positions).  Few blocks: the launches are made whatever the lists hold,
// and an empty block of the exact kernel (256 VGPRs) still waits for a free
// SIMD while the grid chain's kernels occupy the chip.
constexpr unsigned kFixBlocks = 64;
constexpr unsigned kRadBlocks = 256;
// ---- exact re-run of single flagged queries ------------------------------
// A lane of the packed-key query fails its check only through the class of
// its k-th key (a displaced candidate in that class, or two of its first k
// sharing a class).  Its list's k keys all have d2 at most R2 = class_ub(k-th
// key), so the exact top k -- the k smallest (d2, id) -- lie within R2.  One
// wave per such query walks the box tree level by level from the root,
// keeping the boxes whose gap (the walk's own monotone FP64 gap) does not
// exceed R2, gathers the candidates of the buckets it reaches with d2 <= R2,
// sorts them by (d2, id) in LDS and writes the list and the W4 density with
// the exact query's arithmetic.  A query whose frontier or ball outgrows the
// LDS arrays (or, which cannot happen, holds fewer than k) goes to the
// whole-wave exact re-run.
constexpr int kRadFront = 1024;  // frontier nodes of one level
constexpr int kRadCand = 512;    // candidates within R2

template <class P>
__global__ void __launch_bounds__(kWave)
knn_radius_k(const double *qx, const double *qy, const double *qz, const unsigned *qid,
             const double *cxa, const double *cya, const double *cza, const unsigned *cida,
             const Box *boxes, const KHalo *halos, KHalo halo0, const KWave *waves, int k,
             long long *nbr, float *rho, int rho_mode, const P *mass, const unsigned *rad,
             const double *rad_r2, unsigned *whole_count, unsigned *whole_list) {
    __shared__ int s_front[2][kRadFront];
    __shared__ double s_d[kRadCand];
    __shared__ unsigned s_id[kRadCand];
    const int lane = threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned n = rad[0];
    for (unsigned it = blockIdx.x; it < n; it += gridDim.x) {
        const unsigned ent = rad[1 + it];
        const unsigned w = ent / kWave;
        KWave wv;
        if (waves) {
            wv = waves[w];
        } else {
            wv.halo = 0;
            wv.q0 = (long long)w * kWave;
        }
        const KHalo &hd = halos ? halos[wv.halo] : halo0;
        const long long qi = wv.q0 + (long long)(ent % kWave);
        const double p[3] = {qx[qi], qy[qi], qz[qi]};
        const double R2 = rad_r2[it];
        const double *cx = cxa + hd.c0, *cy = cya + hd.c0, *cz = cza + hd.c0;
        const unsigned *cid = cida + hd.c0;
        // the boxes of each level within R2, from the root (node 0 of the
        // top level) down to the leaf buckets
        int cur = 0, nf = 1;
        bool over = false;
        if (lane == 0) s_front[0][0] = 0;
        __syncthreads();
        for (int lev = hd.nlev - 1; lev >= 1 && !over; lev--) {
            const long long nchild = level_count(hd.nb, lev - 1);
            const Box *bl = boxes + hd.loff[lev - 1];
            int cnt = 0;
            for (int c0 = 0; c0 < nf * kFan; c0 += kWave) {
                const int c = c0 + lane;
                bool pass = false;
                long long child = 0;
                if (c < nf * kFan) {
                    child = (long long)s_front[cur][c / kFan] * kFan + c % kFan;
                    if (child < nchild) pass = !(box_gap2(bl[child], p, p) > R2);
                }
                const unsigned long long m = __ballot(pass);
                const int pos = cnt + __popcll(m & below);
                if (pass && pos < kRadFront) s_front[cur ^ 1][pos] = (int)child;
                cnt += __popcll(m);
            }
            over = cnt > kRadFront;
            cur ^= 1;
            nf = cnt;
            __syncthreads();
        }
        // the candidates of those buckets within R2
        int nk = 0;
        if (!over) {
            for (int c0 = 0; c0 < nf * kB; c0 += kWave) {
                const int c = c0 + lane;
                bool in = false;
                double d = 0.0;
                unsigned id = 0u;
                if (c < nf * kB) {
                    const long long e = (long long)s_front[cur][c / kB] * kB + c % kB;
                    if (e < hd.nc) {
                        d = d2_ref(p[0], p[1], p[2], cx[e], cy[e], cz[e]);
                        in = !(d > R2);
                        id = cid[e];
                    }
                }
                const unsigned long long m = __ballot(in);
                const int pos = nk + __popcll(m & below);
                if (in && pos < kRadCand) {
                    s_d[pos] = d;
                    s_id[pos] = id;
                }
                nk += __popcll(m);
            }
            over = nk > kRadCand || nk < k;
        }
        if (over) {
            if (lane == 0) whole_list[atomicAdd(whole_count, 1u)] = w;
            __syncthreads();
            continue;
        }
        // bitonic sort by (d2, id), padded to a power of two
        int N = 1;
        while (N < nk) N <<= 1;
        for (int c = nk + lane; c < N; c += kWave) {
            s_d[c] = __longlong_as_double(0x7ff0000000000000ll);
            s_id[c] = 0xffffffffu;
        }
        __syncthreads();
        for (int size = 2; size <= N; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int c = lane; c < N / 2; c += kWave) {
                    const int lo = 2 * c - (c & (stride - 1)), hi = lo + stride;
                    const double da = s_d[lo], db = s_d[hi];
                    const unsigned ia = s_id[lo], ib = s_id[hi];
                    const bool gt = (da > db) | ((da == db) & (ia > ib));
                    if (gt == ((lo & size) == 0)) {
                        s_d[lo] = db;
                        s_d[hi] = da;
                        s_id[lo] = ib;
                        s_id[hi] = ia;
                    }
                }
                __syncthreads();
            }
        }
        const long long q = qid[qi];
        if (nbr)
            for (int t = lane; t < k; t += kWave)
                nbr[q * k + t] = (long long)s_id[t] - hd.idx_base;
        if (rho && lane == 0) {
            // the exact query's density: h from the k-th distance, the terms
            // in list order
            const float neardist = sqrtf((float)s_d[k - 1]);
            const float h = (float)((double)neardist / 2.0);
            float sum = 0.0f;
            for (int t = 0; t < k; t++) {
                const float r = sqrtf((float)s_d[t]);
                float wt = W4(r, h);
                if (rho_mode == 2) wt = wt * (float)mass[s_id[t]];
                sum += wt;
            }
            rho[q] = sum;
        }
        __syncthreads();
    }
}

// The packed-key query unless the schedule asks for the exact query alone, a
// halo needs more than kMaxKeyBits slot bits, or the debug counters (exact
// query only) are on.
bool use_packed(const Schedule &sc, int max_kb, bool debug) {
    return sc.knn_packed != 0 && !debug && max_kb <= kMaxKeyBits;
}

// Runs the query over nwaves logical waves: the packed-key kernel and the
// exact kernel over the waves it flags, or the exact kernel alone.
template <class P>
hipError_t run_query(bool packed, unsigned nwaves, const double *qx, const double *qy,
                     const double *qz, const unsigned *qid, const double *cx, const double *cy,
                     const double *cz, const unsigned *cid, const Box *boxes, const KHalo *halos,
                     const KHalo &h0, const KWave *waves, int k, long long *nbr, float *rho,
                     int rho_mode, const P *m, int self_seed, unsigned long long *dbg,
                     unsigned xcd_chunk, Workspace &ws, hipStream_t s) {
    if (nwaves == 0) return hipSuccess;
    if (!packed) {
        knn_query_k<P, false><<<nwaves, kWave, 0, s>>>(qx, qy, qz, qid, cx, cy, cz, cid, boxes,
                                                       halos, h0, waves, k, nbr, rho, rho_mode,
                                                       m, self_seed, dbg, nullptr, nullptr,
                                                       xcd_chunk);
        return hipGetLastError();
    }
    // whole waves to re-run (count, then wave ids: the packed kernel's and the
    // radius kernel's), and single flagged queries (count, then wave * 64 +
    // lane) with their radii
    unsigned *fix = ws.alloc_n<unsigned>(1 + (size_t)(1 + kRadLanes) * nwaves);
    unsigned *rad = ws.alloc_n<unsigned>(1 + (size_t)kRadLanes * nwaves);
    double *rad_r2 = ws.alloc_n<double>((size_t)kRadLanes * nwaves);
    if (!fix || !rad || !rad_r2) return hipErrorOutOfMemory;
    PGALF_HIP_TRY(fill_u32(fix, 1, 0u, s));
    PGALF_HIP_TRY(fill_u32(rad, 1, 0u, s));
    static const bool prof_on = [] {
        const char *v = getenv("PGALF_KNN_PROF");
        return v && *v && *v != '0';
    }();
    unsigned long long *prof = nullptr;
    if (prof_on) {
        prof = ws.alloc_n<unsigned long long>(kPrN);
        if (!prof) return hipErrorOutOfMemory;
        PGALF_HIP_TRY(fill_u64(prof, kPrN, 0ull, s));
        knn_packed_k<P, true><<<nwaves, kWave, 0, s>>>(qx, qy, qz, qid, cx, cy, cz, cid, boxes,
                                                       halos, h0, waves, k, nbr, rho, rho_mode,
                                                       m, self_seed, fix, fix + 1, rad, rad_r2,
                                                       prof, xcd_chunk);
    } else {
        timing_mark("knn_packed", s, false);
        knn_packed_k<P, false><<<nwaves, kWave, 0, s>>>(qx, qy, qz, qid, cx, cy, cz, cid,
                                                        boxes, halos, h0, waves, k, nbr, rho,
                                                        rho_mode, m, self_seed, fix, fix + 1,
                                                        rad, rad_r2, nullptr, xcd_chunk);
        timing_mark("knn_packed", s, true);
    }
    const unsigned nfix = nwaves < kFixBlocks ? nwaves : kFixBlocks;
    const unsigned nrad = nwaves < kRadBlocks ? nwaves : kRadBlocks;
    knn_radius_k<P><<<nrad, kWave, 0, s>>>(qx, qy, qz, qid, cx, cy, cz, cid, boxes, halos, h0,
                                           waves, k, nbr, rho, rho_mode, m, rad, rad_r2, fix,
                                           fix + 1);
    knn_query_k<P, true><<<nfix, kWave, 0, s>>>(qx, qy, qz, qid, cx, cy, cz, cid, boxes, halos,
                                                h0, waves, k, nbr, rho, rho_mode, m, self_seed,
                                                nullptr, fix + 1, fix, xcd_chunk);
    if (prof) {
        unsigned long long d[kPrN];
        unsigned nf = 0, nr = 0;
        PGALF_HIP_TRY(hipMemcpyAsync(d, prof, sizeof(d), hipMemcpyDeviceToHost, s));
        PGALF_HIP_TRY(hipMemcpyAsync(&nf, fix, sizeof(nf), hipMemcpyDeviceToHost, s));
        PGALF_HIP_TRY(hipMemcpyAsync(&nr, rad, sizeof(nr), hipMemcpyDeviceToHost, s));
        PGALF_HIP_TRY(hipStreamSynchronize(s));
        const double w = d[kPrWaves] ? (double)d[kPrWaves] : 1.0;
        fprintf(stderr, "[pgalf] packed knn: waves %llu flagged %llu (single queries %u, whole "
                        "waves %u) | kcycles per wave: seed %.1f "
                        "walk %.1f scan %.1f (merge %.1f) tail %.1f total %.1f | per wave: walk "
                        "rounds %.1f scan batches %.1f staged buckets %.1f merged chunks %.1f\n",
                d[kPrWaves], d[kPrFlagged], nr, nf, d[kPrSeed] / w / 1e3, d[kPrWalk] / w / 1e3, d[kPrScan] / w / 1e3,
                d[kPrMerge] / w / 1e3, d[kPrTail] / w / 1e3, d[kPrTotal] / w / 1e3,
                d[kPrRounds] / w, d[kPrBatches] / w, d[kPrSlots] / w, d[kPrChunks] / w);
    }
    return hipGetLastError();
}

}  // namespace

template <class P>
hipError_t knn(const P *x, const P *y, const P *z, const P *m, long long np, const int *type,
               int cand_ptype, int k, long long *nbr, float *rho, int rho_mode, Workspace &ws,
               hipStream_t s, const SortedHook *sorted_hook) {
    if (k < 1 || k > kTop || np < 1) return hipErrorInvalidValue;
    WsScope scope(ws);
    timing_mark("knn_prep", s, false);
    const bool subset = type != nullptr && cand_ptype != PGALF_TYPE_ALL;
    // bounding box of every particle (queries include all candidates)
    unsigned long long *mm = ws.alloc_n<unsigned long long>(6);
    const bool debug = getenv("PGALF_DEBUG") != nullptr;
    unsigned long long *dbg = debug ? ws.alloc_n<unsigned long long>(kDbgN) : nullptr;
    if (!mm || (debug && !dbg)) return hipErrorOutOfMemory;
    if (dbg) PGALF_HIP_TRY(hipMemsetAsync(dbg, 0, kDbgN * sizeof(unsigned long long), s));
    bbox_init_k<<<1, 64, 0, s>>>(mm, 1);
    bbox_k<P><<<256, 256, 0, s>>>(x, y, z, np, mm);

    // queries: all particles, Hilbert order
    unsigned *qk = ws.alloc_n<unsigned>(np), *qv = ws.alloc_n<unsigned>(np);
    unsigned *qk2 = ws.alloc_n<unsigned>(np), *qv2 = ws.alloc_n<unsigned>(np);
    if (!qk || !qv || !qk2 || !qv2) return hipErrorOutOfMemory;
    P *packed = ws.alloc_n<P>(4 * (size_t)np);
    if (!packed) return hipErrorOutOfMemory;
    hilbert_k<P><<<nblk(np, 256), 256, 0, s>>>(x, y, z, nullptr, np, mm, qk, qv, packed);
    PGALF_HIP_TRY(radix_sort_pairs(&qk, &qv, qk2, qv2, np, 30, ws, s));
    double *qx = ws.alloc_n<double>(np), *qy = ws.alloc_n<double>(np),
           *qz = ws.alloc_n<double>(np);
    if (!qx || !qy || !qz) return hipErrorOutOfMemory;
    // every particle a candidate: the leaf boxes come with the gather
    Box *leaf_boxes = nullptr;
    if (!subset) {
        const long long nb = (np + kB - 1) / kB;
        int nlev = 1;
        while (level_count(nb, nlev - 1) > 1 && nlev <= kMaxLevels) nlev++;
        leaf_boxes = ws.alloc_n<Box>((size_t)level_offset(nb, nlev));
        if (!leaf_boxes) return hipErrorOutOfMemory;
        gather_leaf_k<P><<<nblk(nb * kB, 256), 256, 0, s>>>(packed, qv, np, qx, qy, qz, nb,
                                                             leaf_boxes);
    } else {
        gather_packed_k<P><<<nblk(np, 256), 256, 0, s>>>(packed, qv, np, qx, qy, qz);
    }

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
        hilbert_k<P><<<nblk(nc, 256), 256, 0, s>>>(x, y, z, sel, nc, mm, ck, cv, nullptr);
        PGALF_HIP_TRY(radix_sort_pairs(&ck, &cv, ck2, cv2, nc, 30, ws, s));
        gather_sorted_k<P><<<nblk(nc, 256), 256, 0, s>>>(x, y, z, cv, nc, sx, sy, sz);
        cx = sx;
        cy = sy;
        cz = sz;
        cid = cv;
    } else if (np < k) {
        return hipErrorNotReady;
    }

    // 8-ary box tree over kB-candidate buckets
    KHalo h0{};
    long long nb = (nc + kB - 1) / kB;
    if (nb >= (1LL << kRingShift)) return hipErrorInvalidValue;
    h0.c0 = 0;
    h0.nc = nc;
    h0.q0 = 0;
    h0.nq = np;
    h0.idx_base = 0;
    h0.nb = nb;
    h0.kb = key_bits(nb * kB);
    h0.nlev = 1;
    while (level_count(nb, h0.nlev - 1) > 1) h0.nlev++;
    if (h0.nlev > kMaxLevels || h0.nlev > kMaxWalkLevels) return hipErrorInvalidValue;
    for (int l = 0; l < h0.nlev; l++) h0.loff[l] = level_offset(nb, l);
    long long total = level_offset(nb, h0.nlev);
    Box *boxes = leaf_boxes;
    if (!boxes) {
        boxes = ws.alloc_n<Box>((size_t)total);
        if (!boxes) return hipErrorOutOfMemory;
        leaf_boxes_k<<<nblk(nb * kB, 256), 256, 0, s>>>(cx, cy, cz, nc, nb, boxes);
    }
    for (int l = 1; l < h0.nlev; l++) {
        long long nl = level_count(nb, l);
        parent_boxes_k<<<nblk(nl, 256), 256, 0, s>>>(boxes + level_offset(nb, l - 1),
                                                     level_count(nb, l - 1),
                                                     boxes + level_offset(nb, l), nl);
    }
    timing_mark("knn_prep", s, true);
    if (sorted_hook && *sorted_hook) PGALF_HIP_TRY((*sorted_hook)(qx, qy, qz, qv, np));
    timing_mark("knn_query", s, false);
    long long nqb = (np + kWave - 1) / kWave;
    const Schedule sc = schedule_for(np);
    PGALF_HIP_TRY(run_query<P>(use_packed(sc, h0.kb, debug), (unsigned)nqb, qx, qy, qz, qv, cx,
                               cy, cz, cid, boxes, nullptr, h0, nullptr, k, nbr, rho, rho_mode, m,
                               subset ? 0 : 1, dbg, (unsigned)sc.knn_xcd_chunk, ws, s));
    timing_mark("knn_query", s, true);
    if (debug) {
        unsigned long long d[kDbgN];
        PGALF_HIP_TRY(hipMemcpyAsync(d, dbg, sizeof(d), hipMemcpyDeviceToHost, s));
        PGALF_HIP_TRY(hipStreamSynchronize(s));
        double w = d[5] ? (double)d[5] : 1.0;
        fprintf(stderr, "[pgalf] knn: np=%lld nc=%lld buckets=%lld levels=%d | per "
                        "wave: scan batches %.1f, mid-scan merges %.1f, walk rounds %.1f, "
                        "list scans %.1f, end merges %.1f, passing/scanned buckets per group "
                        "%.1f/%.1f | kcycles per wave: seed %.1f, walk %.1f, "
                        "scan %.1f (merges %.1f), total %.1f | accepted per query %.1f, chunk "
                        "merges per wave %.1f\n",
                np, nc, nb, h0.nlev, d[0] / w, d[1] / w, d[2] / w, d[4] / w, d[6] / w,
                d[3] / w / kGroups, d[7] / w / kGroups,
                d[8] / w / 1e3, d[9] / w / 1e3, d[10] / w / 1e3, d[11] / w / 1e3,
                d[12] / w / 1e3, d[13] / (double)(np > 0 ? np : 1), d[14] / w);
#ifdef PGALF_KNN_STATS
        const double nq = (double)(np > 0 ? np : 1);
        fprintf(stderr, "[pgalf] knn stats: accepted per query seed %.1f walk %.1f | wave max "
                        "%.1f (seed %.1f) | group max %.1f | histogram (16s):",
                d[16] / nq, d[17] / nq, d[18] / w, d[19] / w, d[20] / w / kGroups);
        for (int b = 0; b < 32; b++) fprintf(stderr, " %llu", d[24 + b]);
        fprintf(stderr, "\n");
#endif
    }
    return hipGetLastError();
}

// ---- adjoint of the W4 kNN density (no counterpart in the reference) -------
namespace {

// M4 spline of W(r, h) = f(q) / (pi h^3), q = r/h, and f'(q) (nnost.c:536-555).
__device__ inline void m4_f(double q, double &f, double &fp) {
    if (q < 1.0) {
        f = 1.0 - 1.5 * q * q + 0.75 * q * q * q;
        fp = -3.0 * q + 2.25 * q * q;
    } else if (q < 2.0) {
        const double t = 2.0 - q;
        f = 0.25 * t * t * t;
        fp = -0.75 * t * t;
    } else {
        f = 0.0;
        fp = 0.0;
    }
}

// One thread per particle i over its forward list (FP64):
//   rho_i = sum_t m_t W(r_it, h_i), h_i = r_i,(k-1) / 2,
//   dW/dr = f'(q) / (pi h^4), dW/dh = -(3 f + q f') / (pi h^4);
// i's own terms and its neighbours' (opposite sign) are accumulated in FP64.
// A workgroup takes kBwdBlock consecutive particles of the Hilbert order
// (order[]; rank[] is its inverse), so most neighbours lie within
// kBwdHalo sorted positions of the block: their terms go to an LDS window
// (LDS atomics), flushed to acc once per block; the rest go to acc by
// global FP64 atomics (acc = [gx | gy | gz | gm]).
constexpr int kBwdBlock = 256;
constexpr int kBwdHalo = 512;
constexpr int kBwdWin = kBwdBlock + 2 * kBwdHalo;
#ifndef PGALF_KNN_BWD_CHUNK
#define PGALF_KNN_BWD_CHUNK 8
#endif
constexpr int kBwdChunk = PGALF_KNN_BWD_CHUNK;  // neighbours gathered before their adds

template <class P>
__global__ void __launch_bounds__(kBwdBlock)
knn_backward_k(const P *x, const P *y, const P *z, const P *m, long long np,
               const long long *nbr, int k, int weighted, const float *gin, double *acc,
               const unsigned *order, const unsigned *rank, const unsigned *halo_of,
               const HaloSeg *segs, unsigned xcd_chunk) {
    __shared__ double win[4][kBwdWin];
    for (int c = threadIdx.x; c < 4 * kBwdWin; c += kBwdBlock) (&win[0][0])[c] = 0.0;
    __syncthreads();
    // the same XCD-aware order as the query: neighbouring blocks share lists' targets
    const long long s0 = (long long)xcd_block(blockIdx.x, gridDim.x, xcd_chunk) * kBwdBlock;
    const long long wb = s0 - kBwdHalo;  // sorted position of window slot 0
    const long long s = s0 + threadIdx.x;
    // adds v to component c of particle j (sorted position rank[j])
    auto add = [&](long long j, int c, double v) {
        const long long w = (long long)rank[j] - wb;
        if (w >= 0 && w < kBwdWin) unsafeAtomicAdd(&win[c][w], v);
        else unsafeAtomicAdd(&acc[c * np + j], v);
    };
    if (s < np) {
        const long long i = (long long)order[s];
        const double g = (double)gin[i];
        const long long *li = nbr + i * k;
        // batch lists hold indices local to the particle's halo
        const long long base = halo_of ? segs[halo_of[i]].p0 : 0;
        const long long K = li[k - 1] + base;
        const double xi = (double)x[i], yi = (double)y[i], zi = (double)z[i];
        const double kx = xi - (double)x[K], ky = yi - (double)y[K], kz = zi - (double)z[K];
        const double rK = sqrt(kx * kx + ky * ky + kz * kz), hh = 0.5 * rK;
        // g == 0: no terms; k-th neighbour at zero distance: no finite density
        if (g != 0.0 && hh > 0.0) {
            // 1/h, 1/(pi h^3), 1/(pi h^4) once per particle: one square root
            // and one division per neighbour
            const double ih = 1.0 / hh, ih3 = 1.0 / (M_PI * hh * hh * hh), ih4 = ih3 * ih;
            double ox = 0.0, oy = 0.0, oz = 0.0, sh = 0.0;
            // neighbours in chunks: every gather of a chunk is issued before
            // any of its atomics, so the loads overlap instead of waiting
            // behind the previous neighbour's adds
            for (int t0 = 0; t0 < k; t0 += kBwdChunk) {
                long long jj[kBwdChunk];
                double px[kBwdChunk], py[kBwdChunk], pz[kBwdChunk], pm[kBwdChunk];
#pragma unroll
                for (int u = 0; u < kBwdChunk; u++)
                    jj[u] = t0 + u < k ? li[t0 + u] + base : i;
#pragma unroll
                for (int u = 0; u < kBwdChunk; u++) {
                    px[u] = (double)x[jj[u]];
                    py[u] = (double)y[jj[u]];
                    pz[u] = (double)z[jj[u]];
                    pm[u] = weighted ? (double)m[jj[u]] : 1.0;
                }
                double ax[kBwdChunk], ay[kBwdChunk], az[kBwdChunk], am[kBwdChunk];
#pragma unroll
                for (int u = 0; u < kBwdChunk; u++) {
                    ax[u] = ay[u] = az[u] = am[u] = 0.0;
                    if (t0 + u >= k) continue;
                    const double dx = xi - px[u], dy = yi - py[u], dz = zi - pz[u];
                    const double r = sqrt(dx * dx + dy * dy + dz * dz), q = r * ih;
                    double f, fp;
                    m4_f(q, f, fp);
                    const double mt = pm[u];
                    sh -= mt * (3.0 * f + q * fp) * ih4;
                    if (weighted) am[u] = g * f * ih3;
                    if (r > 0.0 && fp != 0.0) {
                        const double c = g * mt * fp * ih4 / r;
                        ox += c * dx;
                        oy += c * dy;
                        oz += c * dz;
                        ax[u] = -c * dx;
                        ay[u] = -c * dy;
                        az[u] = -c * dz;
                    }
                }
#pragma unroll
                for (int u = 0; u < kBwdChunk; u++) {
                    if (t0 + u >= k) continue;
                    if (am[u] != 0.0) add(jj[u], 3, am[u]);
                    if (ax[u] != 0.0 || ay[u] != 0.0 || az[u] != 0.0) {
                        add(jj[u], 0, ax[u]);
                        add(jj[u], 1, ay[u]);
                        add(jj[u], 2, az[u]);
                    }
                }
            }
            const double c = g * sh * 0.5 / rK;
            ox += c * kx;
            oy += c * ky;
            oz += c * kz;
            add(K, 0, -c * kx);
            add(K, 1, -c * ky);
            add(K, 2, -c * kz);
            add(i, 0, ox);
            add(i, 1, oy);
            add(i, 2, oz);
        }
    }
    __syncthreads();
    for (int w = threadIdx.x; w < kBwdWin; w += kBwdBlock) {
        const long long sp = wb + w;
        if (sp < 0 || sp >= np) continue;
        const long long j = (long long)order[sp];
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (win[c][w] != 0.0) unsafeAtomicAdd(&acc[c * np + j], win[c][w]);
    }
}

__global__ void inverse_perm_k(const unsigned *order, long long n, unsigned *rank) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) rank[order[s]] = (unsigned)s;
}

__global__ void grad_finalize_k(const double *acc, long long np, float *gx, float *gy,
                                float *gz, float *gm) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    if (gx) gx[i] = (float)acc[i];
    if (gy) gy[i] = (float)acc[np + i];
    if (gz) gz[i] = (float)acc[2 * np + i];
    if (gm) gm[i] = (float)acc[3 * np + i];
}

}  // namespace

template <class P>
hipError_t knn_density_backward(const P *x, const P *y, const P *z, const P *m, long long np,
                                const long long *nbr, int k, int weighted, const float *gin,
                                double *acc, Workspace &ws, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    if (k < 1 || k > kTop) return hipErrorInvalidValue;
    WsScope scope(ws);
    // Hilbert order of the particles (as the forward's preparation) and its inverse
    unsigned long long *mm = ws.alloc_n<unsigned long long>(6);
    unsigned *hk = ws.alloc_n<unsigned>(np), *hv = ws.alloc_n<unsigned>(np);
    unsigned *hk2 = ws.alloc_n<unsigned>(np), *hv2 = ws.alloc_n<unsigned>(np);
    unsigned *rank = ws.alloc_n<unsigned>(np);
    if (!mm || !hk || !hv || !hk2 || !hv2 || !rank) return hipErrorOutOfMemory;
    bbox_init_k<<<1, 64, 0, s>>>(mm, 1);
    bbox_k<P><<<256, 256, 0, s>>>(x, y, z, np, mm);
    hilbert_k<P><<<nblk(np, 256), 256, 0, s>>>(x, y, z, nullptr, np, mm, hk, hv, nullptr);
    PGALF_HIP_TRY(radix_sort_pairs(&hk, &hv, hk2, hv2, np, 30, ws, s));
    inverse_perm_k<<<nblk(np, 256), 256, 0, s>>>(hv, np, rank);
    knn_backward_k<P><<<nblk(np, kBwdBlock), kBwdBlock, 0, s>>>(
        x, y, z, m, np, nbr, k, weighted, gin, acc, hv, rank, nullptr, nullptr,
        (unsigned)schedule_for(np).knn_xcd_chunk);
    return hipGetLastError();
}

hipError_t grad_finalize(const double *acc, long long np, float *gx, float *gy, float *gz,
                         float *gm, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    grad_finalize_k<<<nblk(np, 256), 256, 0, s>>>(acc, np, gx, gy, gz, gm);
    return hipGetLastError();
}

template hipError_t knn_density_backward<float>(const float *, const float *, const float *,
                                                const float *, long long, const long long *,
                                                int, int, const float *, double *, Workspace &,
                                                hipStream_t);
template hipError_t knn_density_backward<double>(const double *, const double *,
                                                 const double *, const double *, long long,
                                                 const long long *, int, int, const float *,
                                                 double *, Workspace &, hipStream_t);

hipError_t knn_batch_plan(const long long *off, int nhalo, int k, KnnBatchPlan &plan) {
    if (nhalo < 1 || k < 1 || k > kTop || off[0] != 0) return hipErrorInvalidValue;
    KnnBatchPlan p;
    p.sched = schedule_for(off[nhalo]);
    p.nhalo = nhalo;
    p.k = k;
    p.np = off[nhalo];
    if (p.np >= (1LL << 32)) return hipErrorInvalidValue;  // u32 particle ids
    p.hbits = 0;
    while ((1LL << p.hbits) < nhalo) p.hbits++;
    std::vector<KHalo> hs((size_t)nhalo);
    // leaves in padded candidate order
    long long c0 = 0, maxlev = 1;
    for (int h = 0; h < nhalo; h++) {
        const long long n = off[h + 1] - off[h];
        if (n < k) return hipErrorNotReady;  // mapped to PGALF_ERR_SHORT_LIST
        KHalo &hd = hs[h];
        hd.c0 = c0;
        hd.nc = n;
        hd.q0 = c0;
        hd.nq = n;
        hd.idx_base = off[h];
        hd.nb = (n + kB - 1) / kB;
        if (hd.nb >= (1LL << kRingShift)) return hipErrorInvalidValue;
        hd.kb = key_bits(hd.nb * kB);
        if (hd.kb > p.max_kb) p.max_kb = hd.kb;
        hd.nlev = 1;
        while (level_count(hd.nb, hd.nlev - 1) > 1) hd.nlev++;
        if (hd.nlev > kMaxWalkLevels) return hipErrorInvalidValue;
        hd.loff[0] = c0 / kB;
        c0 += hd.nb * kB;
        if (hd.nlev > maxlev) maxlev = hd.nlev;
    }
    p.npad = c0;
    p.nb_total = c0 / kB;
    // internal levels, level-major
    std::vector<long long> cbeg;
    std::vector<unsigned char> ccnt;
    long long b = p.nb_total;
    for (int l = 1; l < maxlev; l++) {
        for (int h = 0; h < nhalo; h++) {
            KHalo &hd = hs[h];
            if (l >= hd.nlev) continue;
            hd.loff[l] = b;
            const long long nl = level_count(hd.nb, l), nchild = level_count(hd.nb, l - 1);
            for (long long j = 0; j < nl; j++) {
                const long long cfirst = kFan * j;
                const long long clast = cfirst + kFan < nchild ? cfirst + kFan : nchild;
                cbeg.push_back(hd.loff[l - 1] + cfirst);
                ccnt.push_back((unsigned char)(clast - cfirst));
            }
            b += nl;
        }
        p.level_begin.push_back(b);
    }
    p.nboxes = b;
    // waves, largest halos first (their waves run longest)
    std::vector<int> order((size_t)nhalo);
    for (int h = 0; h < nhalo; h++) order[h] = h;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return hs[a].nc > hs[c].nc; });
    std::vector<KWave> wv;
    for (int h : order)
        for (long long q = 0; q < hs[h].nc; q += kWave) {
            KWave w;
            w.halo = h;
            w.q0 = hs[h].c0 + q;
            wv.push_back(w);
        }
    p.nwaves = (long long)wv.size();
    hipError_t e = hipMalloc(&p.halos, sizeof(KHalo) * hs.size());
    if (e == hipSuccess) e = hipMalloc(&p.waves, sizeof(KWave) * wv.size());
    const size_t ni = cbeg.size() ? cbeg.size() : 1;
    if (e == hipSuccess) e = hipMalloc(&p.child_begin, sizeof(long long) * ni);
    if (e == hipSuccess) e = hipMalloc(&p.child_cnt, ni);
    if (e == hipSuccess) e = hipMalloc(&p.slot_shift, sizeof(long long) * hs.size());
    if (e == hipSuccess)
        e = hipMemcpy(p.halos, hs.data(), sizeof(KHalo) * hs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        std::vector<long long> shift(hs.size());
        for (size_t h = 0; h < hs.size(); h++) shift[h] = hs[h].c0 - hs[h].idx_base;
        e = hipMemcpy(p.slot_shift, shift.data(), sizeof(long long) * shift.size(),
                      hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = hipMemcpy(p.waves, wv.data(), sizeof(KWave) * wv.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !cbeg.empty())
        e = hipMemcpy(p.child_begin, cbeg.data(), sizeof(long long) * cbeg.size(),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess && !ccnt.empty())
        e = hipMemcpy(p.child_cnt, ccnt.data(), ccnt.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        knn_batch_plan_free(p);
        return e;
    }
    plan = p;
    return hipSuccess;
}

void knn_batch_plan_free(KnnBatchPlan &plan) {
    if (plan.halos) (void)hipFree(plan.halos);
    if (plan.waves) (void)hipFree(plan.waves);
    if (plan.child_begin) (void)hipFree(plan.child_begin);
    if (plan.child_cnt) (void)hipFree(plan.child_cnt);
    if (plan.slot_shift) (void)hipFree(plan.slot_shift);
    plan = KnnBatchPlan();
}

// The batch's particles in (halo, Hilbert within the halo's own box, index)
// order: *order the sorted particle ids, *hsorted their halos, *halo_of the
// halo of every particle.  Two stable sorts: Hilbert key, then halo.
template <class P>
hipError_t batch_order(const P *x, const P *y, const P *z, const BatchDev &b,
                       const KnnBatchPlan &plan, Workspace &ws, hipStream_t s, unsigned **order,
                       unsigned **hsorted, unsigned **halo_of_out) {
    const long long np = plan.np;
    unsigned long long *mm = ws.alloc_n<unsigned long long>(6 * (size_t)plan.nhalo);
    unsigned *hk = ws.alloc_n<unsigned>(np), *hv = ws.alloc_n<unsigned>(np);
    unsigned *hk2 = ws.alloc_n<unsigned>(np), *hv2 = ws.alloc_n<unsigned>(np);
    unsigned *halo_of = ws.alloc_n<unsigned>(np);
    if (!mm || !hk || !hv || !hk2 || !hv2 || !halo_of) return hipErrorOutOfMemory;
    bbox_init_k<<<nblk(6LL * plan.nhalo, 256), 256, 0, s>>>(mm, plan.nhalo);
    bbox_batch_k<P><<<(unsigned)b.nchunks, 256, 0, s>>>(x, y, z, b.chunks, mm);
    hilbert_batch_k<P><<<(unsigned)b.nchunks, 256, 0, s>>>(x, y, z, b.chunks, mm, hk, hv,
                                                          halo_of);
    PGALF_HIP_TRY(radix_sort_pairs(&hk, &hv, hk2, hv2, np, 30, ws, s));
    unsigned *k2 = hk == hk2 ? ws.alloc_n<unsigned>(np) : hk2;
    unsigned *k2t = ws.alloc_n<unsigned>(np), *v2t = hv == hv2 ? ws.alloc_n<unsigned>(np) : hv2;
    if (!k2 || !k2t || !v2t) return hipErrorOutOfMemory;
    halo_keys_k<<<nblk(np, 256), 256, 0, s>>>(hv, halo_of, np, k2);
    if (plan.hbits > 0) PGALF_HIP_TRY(radix_sort_pairs(&k2, &hv, k2t, v2t, np, plan.hbits, ws, s));
    *order = hv;
    *hsorted = k2;
    *halo_of_out = halo_of;
    return hipGetLastError();
}

template <class P>
hipError_t knn_batch(const P *x, const P *y, const P *z, const P *m, const BatchDev &b,
                     const KnnBatchPlan &plan, long long *nbr, float *rho, int rho_mode,
                     Workspace &ws, hipStream_t s, const SortedHook *sorted_hook) {
    const long long np = plan.np, npad = plan.npad;
    if (np != b.np || plan.nhalo != b.nhalo) return hipErrorInvalidValue;
    WsScope scope(ws);
    timing_mark("knn_prep", s, false);
    const KHalo *halos = static_cast<const KHalo *>(plan.halos);
    unsigned *hv = nullptr, *k2 = nullptr, *halo_of = nullptr;
    PGALF_HIP_TRY(batch_order<P>(x, y, z, b, plan, ws, s, &hv, &k2, &halo_of));
    double *sx = ws.alloc_n<double>(npad), *sy = ws.alloc_n<double>(npad),
           *sz = ws.alloc_n<double>(npad);
    unsigned *sid = ws.alloc_n<unsigned>(npad);
    Box *boxes = ws.alloc_n<Box>((size_t)plan.nboxes);
    if (!sx || !sy || !sz || !sid || !boxes) return hipErrorOutOfMemory;
    gather_padded_k<P><<<nblk(np, 256), 256, 0, s>>>(x, y, z, hv, k2, halos, np, sx, sy, sz, sid);
    pad_slots_k<<<nblk((long long)plan.nhalo * kB, 256), 256, 0, s>>>(halos, plan.nhalo, sx, sy,
                                                                      sz, sid);
    leaf_boxes_k<<<nblk(plan.nb_total * kB, 256), 256, 0, s>>>(sx, sy, sz, npad, plan.nb_total,
                                                          boxes);
    long long lb = plan.nb_total;
    for (long long le : plan.level_begin) {
        node_boxes_k<<<nblk(le - lb, 256), 256, 0, s>>>(boxes, lb, le, plan.nb_total,
                                                        plan.child_begin, plan.child_cnt);
        lb = le;
    }
    timing_mark("knn_prep", s, true);
    if (sorted_hook && *sorted_hook) PGALF_HIP_TRY((*sorted_hook)(sx, sy, sz, sid, npad));
    timing_mark("knn_query", s, false);
    KHalo h0{};
    PGALF_HIP_TRY(run_query<P>(use_packed(plan.sched, plan.max_kb, false), (unsigned)plan.nwaves,
                               sx, sy, sz, sid, sx, sy, sz, sid, boxes, halos, h0,
                               static_cast<const KWave *>(plan.waves), plan.k, nbr, rho,
                               rho_mode, m, 1, nullptr, (unsigned)plan.sched.knn_xcd_chunk, ws,
                               s));
    timing_mark("knn_query", s, true);
    return hipGetLastError();
}

template <class P>
hipError_t knn_batch_backward(const P *x, const P *y, const P *z, const P *m, const BatchDev &b,
                              const KnnBatchPlan &plan, const long long *nbr, int weighted,
                              const float *gin, double *acc, Workspace &ws, hipStream_t s) {
    const long long np = plan.np;
    if (np != b.np || plan.nhalo != b.nhalo) return hipErrorInvalidValue;
    WsScope scope(ws);
    unsigned *order = nullptr, *hsorted = nullptr, *halo_of = nullptr;
    PGALF_HIP_TRY(batch_order<P>(x, y, z, b, plan, ws, s, &order, &hsorted, &halo_of));
    unsigned *rank = ws.alloc_n<unsigned>(np);
    if (!rank) return hipErrorOutOfMemory;
    inverse_perm_k<<<nblk(np, 256), 256, 0, s>>>(order, np, rank);
    knn_backward_k<P><<<nblk(np, kBwdBlock), kBwdBlock, 0, s>>>(
        x, y, z, m, np, nbr, plan.k, weighted, gin, acc, order, rank, halo_of, b.segs,
        (unsigned)plan.sched.knn_xcd_chunk);
    return hipGetLastError();
}

template hipError_t knn_batch_backward<float>(const float *, const float *, const float *,
                                              const float *, const BatchDev &,
                                              const KnnBatchPlan &, const long long *, int,
                                              const float *, double *, Workspace &, hipStream_t);

template hipError_t knn_batch<float>(const float *, const float *, const float *, const float *,
                                     const BatchDev &, const KnnBatchPlan &, long long *,
                                     float *, int, Workspace &, hipStream_t, const SortedHook *);
template hipError_t knn_batch<double>(const double *, const double *, const double *,
                                      const double *, const BatchDev &, const KnnBatchPlan &,
                                      long long *, float *, int, Workspace &, hipStream_t,
                                      const SortedHook *);

template hipError_t knn<float>(const float *, const float *, const float *, const float *,
                               long long, const int *, int, int, long long *, float *, int,
                               Workspace &, hipStream_t, const SortedHook *);
template hipError_t knn<double>(const double *, const double *, const double *,
                                const double *, long long, const int *, int, int, long long *,
                                float *, int, Workspace &, hipStream_t, const SortedHook *);

}  // namespace pgalf
