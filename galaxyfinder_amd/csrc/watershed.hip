// Watershed of subhalo_den on the GPU: FindCoreDensity
// (NewGalFinder/subhaloden.mod6.c:1167-1475), the shell thresholds
// (:2995-3024) and GetShellParticlesPeaks + SaveRemainingParticles2LastShell
// (:572-678).  Every step is a breadth-first sweep over the directed
// neighbour graph (particle u -> the k entries of its list) through the
// particles denser than a threshold; outputs are integer lists, counts and
// FP32 values formed by the reference's own operations, bit-identical to it.
//
// Sweeps.  The reference's sweep is a FIFO queue; its order is kept exactly:
// the queue is processed two particles (2 x 32 list slots, one wavefront) at
// a time, and a particle reached from both in one step belongs to the first
// (the lowest lane).  The visited set is an LDS hash (one wavefront per
// sweep, up to kWaveQ particles); a sweep that outgrows it is re-run
// level-synchronously over the whole device (candidates of one level ranked
// by (parent position, slot) with an atomicMin claim, then compacted in that
// order by a scan -- the same FIFO order).
//
// FindCoreDensity.  One wavefront per core runs the reference's threshold
// search (halving [minden, den[peak]] while the relative width exceeds
// COREDENRESOLUTION = 1e-3, params.h:12), each trial a sweep that stops at
// the first other peak, then the final sweep at the core density; its
// particles are the core's members in sweep order.  The final sweep cannot
// meet a peak unless the search never found a threshold free of other peaks
// (the core is then dropped).  The reference then leaves the sweep's
// particles marked visited for every later core (it clears the marks of
// `ncontact` particles after setting ncontact = 0, :1301-1310); here such a
// core ends the pass: its particles go into a `stale` bitmap that every later
// sweep treats as visited, and the cores after it are run again.  The
// reference runs the cores on OpenMP threads with per-thread marks; the
// result here is the one-thread (schedule-free) one, and a particle in the
// members of several cores takes the last core's id, as one thread leaves it.
//
// Shells.  Per threshold, the cores above it that no earlier sweep of the
// threshold reached root sweeps in core order.  Sweeps of a batch of
// candidate roots run speculatively in parallel (each is independent of the
// others: the reference clears the marks before every sweep); one workgroup
// then walks the batch in core order, skips the candidates an earlier root
// reached, and applies each root: its peaks' cores enclosed and listed, its
// particles that are neither shell nor core particles listed as the new
// shell.  A root whose sweep outgrew the LDS is swept device-wide and
// applied by scans.
#include <algorithm>
#include <atomic>
#include <system_error>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "pgalf_internal.h"

namespace pgalf {
namespace {

constexpr unsigned char kWpPeak = 0x01, kWpVisited = 0x02, kWpCore = 0x04, kWpShell = 0x08,
                        kWpBound = 0x10, kWpRemaining = 0x20;
constexpr int kNotMember = -9999;  // NOT_HALO_MEMBER, defs.h:29
constexpr int kTypeStar = 1;       // TYPE_STAR, tree.h:47
constexpr double kCoreDenResolution = 1.e-3;  // COREDENRESOLUTION, params.h:12
constexpr double kDenFloor = 1.0;             // DENFLOOR, params.h:109
constexpr int kMaxSearch = 4096;  // threshold-search steps before giving up

constexpr int kWaveQ = 512;                 // particles one wavefront sweep holds
constexpr int kWaveH = 1024;                // hash slots (load <= 1/2)
constexpr int kWaveLimit = kWaveQ - 64;     // a step starts only at or below this
constexpr int kHashBits = 10;
static_assert((1 << kHashBits) == kWaveH, "hash size");

enum Mode { kTrial = 0, kFinal = 1, kShell = 2 };
// kBig: too large for the LDS sweep; kMid: swept by a workgroup; kGiant: left
// to the device-wide sweep.
// kHuge: a shell root swept device-wide with others (multi_sweep).
enum : unsigned char { kOk = 0, kQuirk = 1, kBig = 2, kSearchFail = 3, kMid = 4, kGiant = 5,
                       kHuge = 6 };
enum : int { kSweepHit = -1, kSweepBig = -2 };

inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ bool stale_bit(const unsigned *stale, int v) {
    return stale && ((stale[v >> 5] >> (v & 31)) & 1u);
}

// Candidate test of one list entry v at threshold t: a particle to push
// (cand) or an unvisited other peak (hit).  Peaks are pushed only in shell
// sweeps; in core sweeps the root is the one visited peak.
__device__ __forceinline__ void entry_test(int mode, int v, int root, float t, const float *den,
                                           const unsigned char *pk, const unsigned *stale,
                                           bool &cand, bool &hit) {
    // both flags are computed as values and stored once each: a store to
    // one flag or the other would keep them in scratch memory
    bool c = false, h = false;
    if (den[v] > t) {
        if (mode == kShell) {
            c = true;
        } else if (!stale_bit(stale, v)) {
            const bool peak = pk[v] != 0;
            c = !peak;
            h = peak && v != root;
        }
    }
    cand = c;
    hit = h;
}

struct WaveLds {
    int hash[kWaveH];  // visited particle ids, -1 empty
    int own[kWaveH];   // lowest lane claiming a new entry this step; -1 once visited
    int q[kWaveQ];     // the sweep's queue
};

__device__ __forceinline__ int hash_slot(int *hash, int v) {
    unsigned h = ((unsigned)v * 2654435761u) >> (32 - kHashBits);
    for (;;) {
        const int old = atomicCAS(&hash[h], -1, v);
        if (old == -1 || old == v) return (int)h;
        h = (h + 1) & (kWaveH - 1);
    }
}

// One FIFO sweep by a 64-lane workgroup from root at threshold t.  Returns
// the particle count (queue L.q[0 .. n)), kSweepHit when a trial sweep meets
// another peak, kSweepBig when the sweep outgrows the LDS.  A final sweep
// that meets a peak skips the rest of that particle's list (the
// reference's `break` out of the slot loop) and sets *fhit.
__device__ int wave_sweep(WaveLds &L, int mode, int root, float t, const float *den,
                          const long long *nbr, int k, const unsigned char *pk,
                          const unsigned *stale, bool *fhit) {
    const int lane = threadIdx.x;
    for (int i = lane; i < kWaveH; i += 64) {
        L.hash[i] = -1;
        L.own[i] = INT_MAX;
    }
    __syncthreads();
    if (lane == 0) {
        L.q[0] = root;
        L.own[hash_slot(L.hash, root)] = -1;
    }
    __syncthreads();
    int n = 1, head = 0;
    bool anyhit = false;
    const int f = lane >> 5, s = lane & 31;
    while (head < n) {
        if (n > kWaveLimit) return kSweepBig;
        const int nodes = n - head < 2 ? n - head : 2;
        int v = -1;
        bool cand = false, hit = false;
        if (f < nodes && s < k) {
            const int u = L.q[head + f];
            v = (int)nbr[(long long)u * k + s];
            entry_test(mode, v, root, t, den, pk, stale, cand, hit);
        }
        const unsigned long long hits = __ballot(hit);
        if (hits) {
            if (mode == kTrial) return kSweepHit;
            anyhit = true;
            const unsigned gh = (unsigned)(hits >> (32 * f));
            if (gh && s >= __builtin_ctz(gh)) cand = false;
        }
        int e = -1;
        if (cand) {
            e = hash_slot(L.hash, v);
            atomicMin(&L.own[e], lane);
        }
        __syncthreads();
        const bool push = cand && L.own[e] == lane;
        const unsigned long long pm = __ballot(push);
        if (push) {
            L.q[n + __popcll(pm & ((1ull << lane) - 1ull))] = v;
            L.own[e] = -1;
        }
        n += __popcll(pm);
        head += nodes;
        __syncthreads();
    }
    if (fhit) *fhit = anyhit;
    return n;
}

__device__ __forceinline__ float core_trial(float up, float dn) {
    return (float)(0.5 * (double)(up + dn));
}
__device__ __host__ __forceinline__ bool search_more(float up, float dn, float t) {
    return fabs((double)((up - dn) / t)) > kCoreDenResolution;
}

// ---- set-up ---------------------------------------------------------------

__global__ void ws_init_k(long long np, unsigned char *flag, int *haloid, int *core_of) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    flag[i] = 0;
    if (haloid) haloid[i] = kNotMember;
    core_of[i] = INT_MAX;
}

__global__ void ws_peaks_k(const int *peak, int numcore, long long np, unsigned char *flag,
                           int *core_of, unsigned *err) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numcore) return;
    const int p = peak[c];
    if (p < 0 || p >= np) {
        atomicOr(err, DEV_ERR_BAD_INDEX);
        return;
    }
    flag[p] = kWpPeak;
    atomicMin(&core_of[p], c);
}

// core_of[p]: the first core whose peak is particle p (INT_MAX for none).
__global__ void core_of_k(const int *peak, int numcore, long long np, int *core_of,
                          unsigned *err) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numcore) return;
    const int p = peak[c];
    if (p < 0 || p >= np) {
        atomicOr(err, DEV_ERR_BAD_INDEX);
        return;
    }
    atomicMin(&core_of[p], c);
}

// Peak flags as one byte per particle (read by the sweeps) from the flag bits.
__global__ void ws_peakbyte_k(const unsigned char *flag, long long np, unsigned char *pk) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) pk[i] = flag[i] & kWpPeak;
}

// Order-preserving unsigned key of a float; min over keys -> min over floats.
__device__ __forceinline__ unsigned ord_key(float v) {
    const unsigned b = __float_as_uint(v);
    return (b >> 31) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord_val(unsigned k) {
    return __uint_as_float((k >> 31) ? (k & 0x7fffffffu) : ~k);
}

__global__ void ws_minmax_k(const float *a, long long n, unsigned *kmin, unsigned *kmax) {
    unsigned lo = 0xffffffffu, hi = 0u;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const unsigned kk = ord_key(a[i]);
        lo = kk < lo ? kk : lo;
        hi = kk > hi ? kk : hi;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned a2 = __shfl_down(lo, o), b2 = __shfl_down(hi, o);
        lo = a2 < lo ? a2 : lo;
        hi = b2 > hi ? b2 : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        if (kmin) atomicMin(kmin, lo);
        if (kmax) atomicMax(kmax, hi);
    }
}

__global__ void ws_minmax_init_k(unsigned *kk) {
    kk[0] = 0xffffffffu;
    kk[1] = 0u;
}

// ---- FindCoreDensity, one wavefront per core -----------------------------

struct CoreArgs {
    const float *den;
    const long long *nbr;
    int k;
    const unsigned char *pk;
    const unsigned *stale;
    const int *type;
    const double *mass;
    const int *peak;
    int numcore;
    const unsigned *kmin;  // ordered key of the smallest density
    float *coredensity;
    int *nummem, *numstar;
    float *starmass;
    unsigned char *status;
    int *memb;   // [core][kWaveQ] members (or the dropped sweep's particles)
    int *count;  // [core]
};

__global__ __launch_bounds__(64) void core_wave_k(CoreArgs a, int c0) {
    __shared__ WaveLds L;
    const int c = c0 + (int)blockIdx.x;
    if (c >= a.numcore) return;
    const int lane = threadIdx.x;
    const int root = a.peak[c];
    const float gmin = ord_val(*a.kmin) < 1.e23f ? ord_val(*a.kmin) : 1.e23f;
    const float minden = (float)(kDenFloor > (double)gmin ? kDenFloor : (double)gmin);
    float up = a.den[root], dn = minden, t;
    unsigned char st = kOk;
    int steps = 0;
    do {
        t = core_trial(up, dn);
        const int r = wave_sweep(L, kTrial, root, t, a.den, a.nbr, a.k, a.pk, a.stale, nullptr);
        if (r == kSweepBig) {
            st = kBig;
            break;
        }
        if (r == kSweepHit) dn = t;
        else up = t;
        if (++steps > kMaxSearch) {
            st = kSearchFail;
            break;
        }
    } while (search_more(up, dn, t));
    int n = 0;
    bool fhit = false;
    if (st == kOk) {
        n = wave_sweep(L, kFinal, root, up, a.den, a.nbr, a.k, a.pk, a.stale, &fhit);
        if (n == kSweepBig) st = kBig;
        else if (fhit) st = kQuirk;
    }
    if (st == kBig || st == kSearchFail) {
        if (lane == 0) a.status[c] = st;
        return;
    }
    int *out = a.memb + (long long)c * kWaveQ;
    for (int j = lane; j < n; j += 64) out[j] = L.q[j];
    if (st == kQuirk) {
        if (lane == 0) {
            a.status[c] = st;
            a.count[c] = n;
            a.coredensity[c] = up;
            a.nummem[c] = 0;
            a.numstar[c] = 0;
            a.starmass[c] = 0.0f;
        }
        return;
    }
    // stars' masses in sweep order, summed in FP32 one by one (the reference's
    // `corestarmass += bp[jj].mass` with a float accumulator)
    static_assert(sizeof(double) * kWaveQ <= sizeof(int) * kWaveH, "staging fits");
    double *ms = reinterpret_cast<double *>(L.hash);  // the hash is free after the sweep
    int *isstar = L.own;
    __syncthreads();
    for (int j = lane; j < n; j += 64) {
        const int v = L.q[j];
        const bool s1 = a.type[v] == kTypeStar;
        ms[j] = s1 ? a.mass[v] : 0.0;
        isstar[j] = s1;
    }
    __syncthreads();
    if (lane == 0) {
        float acc = 0.0f;
        int ns = 0;
        for (int j = 0; j < n; j++)
            if (isstar[j]) {
                ns++;
                acc = (float)((double)acc + ms[j]);
            }
        a.status[c] = kOk;
        a.count[c] = n;
        a.coredensity[c] = up;
        a.nummem[c] = n;
        a.numstar[c] = ns;
        a.starmass[c] = acc;
    }
}

// Members of the final cores: CORE and BOUND set (REMAINING cleared), id of
// the last core (the largest index) whose members hold the particle.
__global__ void core_apply_k(const unsigned char *status, const int *count, const int *memb,
                             int c0, int c1, unsigned char *flag, int *haloid) {
    const int c = c0 + (int)blockIdx.x;
    if (c >= c1 || status[c] != kOk) return;
    const int n = count[c];
    const int *m = memb + (long long)c * kWaveQ;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int v = m[j];
        flag[v] = (unsigned char)((flag[v] | kWpCore | kWpBound) & ~kWpRemaining);
        atomicMax(&haloid[v], c);
    }
}

__global__ void list_apply_k(const int *list, long long n, int c, unsigned char *flag,
                             int *haloid) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int v = list[j];
    flag[v] = (unsigned char)((flag[v] | kWpCore | kWpBound) & ~kWpRemaining);
    atomicMax(&haloid[v], c);
}

__global__ void stale_mark_k(const int *list, long long n, unsigned *stale) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int v = list[j];
    atomicOr(&stale[v >> 5], 1u << (v & 31));
}

// Stars and their FP32 mass sum over a device-wide sweep's list (one block:
// the masses are staged in order, one lane sums them).
__global__ __launch_bounds__(256) void star_sum_k(const int *list, long long n, const int *type,
                                                  const double *mass, int c, int *nummem,
                                                  int *numstar, float *starmass) {
    __shared__ double ms[1024];
    __shared__ unsigned char is[1024];
    float acc = 0.0f;
    int ns = 0;
    for (long long b = 0; b < n; b += 1024) {
        const int m = (int)(n - b < 1024 ? n - b : 1024);
        for (int j = threadIdx.x; j < m; j += blockDim.x) {
            const int v = list[b + j];
            const bool s1 = type[v] == kTypeStar;
            ms[j] = s1 ? mass[v] : 0.0;
            is[j] = s1;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int j = 0; j < m; j++)
                if (is[j]) {
                    ns++;
                    acc = (float)((double)acc + ms[j]);
                }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        nummem[c] = (int)n;
        numstar[c] = ns;
        starmass[c] = acc;
    }
}

__global__ void core_scalar_k(int c, float cd, int nm, int *nummem, int *numstar,
                              float *starmass, float *coredensity, unsigned char *status,
                              unsigned char st) {
    coredensity[c] = cd;
    if (nm >= 0) {
        nummem[c] = nm;
        numstar[c] = 0;
        starmass[c] = 0.0f;
    }
    status[c] = st;
}

template <class T>
hipError_t to_host(T *h, const T *d, size_t n, hipStream_t s) {
    PGALF_HIP_TRY(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}

// ---- device-wide level-synchronous sweep ----------------------------------

struct LevelArgs {
    const float *den;
    const long long *nbr;
    int k;
    const unsigned char *pk;
    const unsigned *stale;
    int mode, root;
    float t;
    int *bq;           // queue, capacity np
    unsigned *vb;      // visited bitmap of this sweep
    int *claim;        // per particle: lowest candidate of this level, INT_MAX idle
    unsigned *cf;      // per candidate: flag, then position
    unsigned *hitflag;
};

__global__ void lvl_cand_k(LevelArgs a, long long lo, long long nf) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long f = i >> 5;
    const int s = (int)(i & 31);
    bool cand = false, hit = false;
    int v = -1;
    if (f < nf && s < a.k) {
        const int u = a.bq[lo + f];
        v = (int)a.nbr[(long long)u * a.k + s];
        entry_test(a.mode, v, a.root, a.t, a.den, a.pk, nullptr, cand, hit);
        if (a.mode != kShell && stale_bit(a.stale, v)) cand = hit = false;
        if (cand && ((a.vb[v >> 5] >> (v & 31)) & 1u)) cand = false;
    }
    const unsigned long long hits = __ballot(hit);
    if (hits) {
        const unsigned gh = (unsigned)(hits >> (32 * ((threadIdx.x >> 5) & 1)));
        if (gh && s >= __builtin_ctz(gh)) cand = false;
        if (hit) atomicOr(a.hitflag, 1u);
    }
    if (f < nf) {
        if (cand) atomicMin(&a.claim[v], (int)i);
        a.cf[i] = cand ? 1u : 0u;
    }
}

// win[i]: candidate i is the first claim of its particle in this level.
__global__ void lvl_win_k(LevelArgs a, long long lo, long long nf, unsigned *win) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if ((i >> 5) >= nf) return;
    unsigned w = 0u;
    if (a.cf[i]) {
        const int u = a.bq[lo + (i >> 5)];
        const int v = (int)a.nbr[(long long)u * a.k + (i & 31)];
        w = a.claim[v] == (int)i ? 1u : 0u;
    }
    win[i] = w;
}

__global__ void lvl_push_k(LevelArgs a, long long lo, long long nf, const unsigned *flags,
                           long long hi) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if ((i >> 5) >= nf) return;
    const int u = a.bq[lo + (i >> 5)];
    const int s = (int)(i & 31);
    if (s >= a.k) return;
    const int v = (int)a.nbr[(long long)u * a.k + s];
    if (flags[i]) {
        a.bq[hi + a.cf[i]] = v;
        atomicOr(&a.vb[v >> 5], 1u << (v & 31));
    }
    if (a.claim[v] == (int)i) a.claim[v] = INT_MAX;
}

__global__ void lvl_root_k(int *bq, unsigned *vb, int root, unsigned *hitflag) {
    bq[0] = root;
    atomicOr(&vb[root >> 5], 1u << (root & 31));
    *hitflag = 0u;
}

__global__ void bits_clear_k(const int *list, long long n, unsigned *vb) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) {
        const int v = list[j];
        atomicAnd(&vb[v >> 5], ~(1u << (v & 31)));
    }
}

// Buffers of the device-wide sweep (allocated once per entry point call).
struct LevelBufs {
    int *bq = nullptr;
    unsigned *vb = nullptr;
    int *claim = nullptr;
    unsigned *cf = nullptr, *flags = nullptr, *small = nullptr;  // small[0] hit, [1] total
    long long np = 0;
};

hipError_t level_bufs(LevelBufs &b, long long np, int k, Workspace &ws, hipStream_t s) {
    b.np = np;
    b.bq = ws.alloc_n<int>(np);
    b.vb = ws.alloc_n<unsigned>((np + 31) / 32);
    b.claim = ws.alloc_n<int>(np);
    b.cf = ws.alloc_n<unsigned>((size_t)np * 32);
    b.flags = ws.alloc_n<unsigned>((size_t)np * 32);
    b.small = ws.alloc_n<unsigned>(4);
    if (!b.bq || !b.vb || !b.claim || !b.cf || !b.flags || !b.small)
        return hipErrorOutOfMemory;
    (void)k;
    PGALF_HIP_TRY(fill_u32(b.vb, (np + 31) / 32, 0u, s));
    return fill_u32(reinterpret_cast<unsigned *>(b.claim), np, 0x7fffffffu, s);
}

// One device-wide FIFO sweep; the queue is b.bq[0 .. *n).  *hit: a peak
// was met (a trial sweep stops at that level).  The visited bits are
// cleared again before returning.
hipError_t level_sweep(LevelBufs &b, int mode, int root, float t, const float *den,
                       const long long *nbr, int k, const unsigned char *pk,
                       const unsigned *stale, long long *n_out, bool *hit_out, Workspace &ws,
                       hipStream_t s) {
    LevelArgs a;
    a.den = den;
    a.nbr = nbr;
    a.k = k;
    a.pk = pk;
    a.stale = stale;
    a.mode = mode;
    a.root = root;
    a.t = t;
    a.bq = b.bq;
    a.vb = b.vb;
    a.claim = b.claim;
    a.hitflag = b.small;
    lvl_root_k<<<1, 1, 0, s>>>(b.bq, b.vb, root, b.small);
    long long lo = 0, hi = 1;
    bool hit = false;
    unsigned hv[2];
    while (lo < hi) {
        const long long nf = hi - lo, nc = nf * 32;
        a.cf = b.cf;
        const unsigned g = nblk(nc, 256);
        lvl_cand_k<<<g, 256, 0, s>>>(a, lo, nf);
        lvl_win_k<<<g, 256, 0, s>>>(a, lo, nf, b.flags);
        // b.cf receives the winners' positions
        {
            WsScope sc(ws);
            PGALF_HIP_TRY(exclusive_scan_u32(b.flags, b.cf, nc, b.small + 1, ws, s));
        }
        lvl_push_k<<<g, 256, 0, s>>>(a, lo, nf, b.flags, hi);
        PGALF_HIP_TRY(hipMemcpyAsync(hv, b.small, sizeof(unsigned) * 2, hipMemcpyDeviceToHost, s));
        PGALF_HIP_TRY(hipStreamSynchronize(s));
        lo = hi;
        hi += hv[1];
        if (hv[0]) {
            hit = true;
            if (mode == kTrial) break;
        }
    }
    bits_clear_k<<<nblk(hi, 256), 256, 0, s>>>(b.bq, hi, b.vb);
    *n_out = hi;
    *hit_out = hit;
    return hipGetLastError();
}

// ---- workgroup sweeps with a global visited bitmap ------------------------
// The middle tier: one 1024-thread workgroup per sweep, 32 queue entries per
// step (one list slot per thread), the visited bits in a per-workgroup
// bitmap of the pool and the queue in global memory (cap entries).  Many
// sweeps run at once; a sweep that outgrows its queue is left to the
// device-wide tier.

constexpr int kBlkT = 1024;
constexpr int kBlkH = 2048;  // per-step claim hash (<= 1024 candidates)

struct BlockLds {
    int key[kBlkH];
    int own[kBlkH];
    int wtot[kBlkT / 64];
    int hitflag;
};

struct Pool {
    unsigned *bits;  // [slots][nwords], all zero between sweeps
    int *q;          // [slots][cap]
    long long nwords, cap;
    int slots;
};

__device__ __forceinline__ int hash_slot2(int *key, int v) {
    unsigned h = ((unsigned)v * 2654435761u) >> (32 - 11);
    for (;;) {
        const int old = atomicCAS(&key[h], -1, v);
        if (old == -1 || old == v) return (int)h;
        h = (h + 1) & (kBlkH - 1);
    }
}

// Block-wide exclusive prefix of a flag over kBlkT threads.
__device__ __forceinline__ int block_prefix_1k(bool fl, int *wtot, int &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long m = __ballot(fl);
    if (lane == 0) wtot[w] = __popcll(m);
    __syncthreads();
    int off = 0;
    total = 0;
#pragma unroll
    for (int j = 0; j < kBlkT / 64; j++) {
        off += j < w ? wtot[j] : 0;
        total += wtot[j];
    }
    __syncthreads();
    return off + __popcll(m & ((1ull << lane) - 1ull));
}

// One FIFO sweep by the workgroup (same contract as wave_sweep; the queue is
// q[0 .. n), n <= cap).  The visited bits are cleared before returning.
__device__ long long block_sweep(BlockLds &L, int mode, int root, float t, const float *den,
                                 const long long *nbr, int k, const unsigned char *pk,
                                 const unsigned *stale, unsigned *vb, int *q, long long cap,
                                 bool *fhit) {
    const int tid = threadIdx.x;
    for (int i = tid; i < kBlkH; i += kBlkT) {
        L.key[i] = -1;
        L.own[i] = INT_MAX;
    }
    if (tid == 0) {
        q[0] = root;
        atomicOr(&vb[root >> 5], 1u << (root & 31));
        L.hitflag = 0;
    }
    __syncthreads();
    long long n = 1, head = 0, ret;
    bool anyhit = false;
    const int f = tid >> 5, s = tid & 31;
    for (;;) {
        if (head >= n) {
            ret = n;
            break;
        }
        if (n > cap - kBlkT) {
            ret = kSweepBig;
            break;
        }
        const long long nodes = n - head < 32 ? n - head : 32;
        int v = -1;
        bool cand = false, hit = false;
        if (f < nodes && s < k) {
            const int u = q[head + f];
            v = (int)nbr[(long long)u * k + s];
            entry_test(mode, v, root, t, den, pk, nullptr, cand, hit);
            // visited and stale bits may be set by atomics of other waves of
            // this workgroup: read them past L1
            if (mode != kShell && stale && (cand || hit) &&
                ((__hip_atomic_load(&stale[v >> 5], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) >> (v & 31)) & 1u))
                cand = hit = false;
            if (cand && ((__hip_atomic_load(&vb[v >> 5], __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT) >>
                          (v & 31)) & 1u))
                cand = false;
        }
        const unsigned long long hits = __ballot(hit);
        if (hits) {
            const unsigned gh = (unsigned)(hits >> (32 * (f & 1)));
            if (gh && s >= __builtin_ctz(gh)) cand = false;
            if ((tid & 63) == 0) L.hitflag = 1;
        }
        int e = -1;
        if (cand) {
            e = hash_slot2(L.key, v);
            atomicMin(&L.own[e], tid);
        }
        __syncthreads();
        const bool push = cand && L.own[e] == tid;
        int tot;
        const int pos = block_prefix_1k(push, L.wtot, tot);
        if (push) {
            q[n + pos] = v;
            atomicOr(&vb[v >> 5], 1u << (v & 31));
        }
        if (cand) {
            L.key[e] = -1;
            L.own[e] = INT_MAX;
        }
        n += tot;
        head += nodes;
        __syncthreads();
        if (L.hitflag) {
            anyhit = true;
            if (mode == kTrial) {
                ret = kSweepHit;
                break;
            }
        }
    }
    __syncthreads();
    for (long long j = tid; j < n; j += kBlkT) {
        const int v = q[j];
        atomicAnd(&vb[v >> 5], ~(1u << (v & 31)));
    }
    if (fhit) *fhit = anyhit;
    __syncthreads();
    return ret;
}

// FP32 sum of the stars' masses of q[0 .. n) in order (staged per 1024).
__device__ void block_star_sum(BlockLds &L, const int *q, long long n, const int *type,
                               const double *mass, int *ns_out, float *acc_out) {
    double *ms = reinterpret_cast<double *>(L.key);
    unsigned char *is = reinterpret_cast<unsigned char *>(L.own);
    float acc = 0.0f;
    int ns = 0;
    for (long long b = 0; b < n; b += kBlkT) {
        const int m = (int)(n - b < kBlkT ? n - b : kBlkT);
        if ((int)threadIdx.x < m) {
            const int v = q[b + threadIdx.x];
            const bool s1 = type[v] == kTypeStar;
            ms[threadIdx.x] = s1 ? mass[v] : 0.0;
            is[threadIdx.x] = s1;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int j = 0; j < m; j++)
                if (is[j]) {
                    ns++;
                    acc = (float)((double)acc + ms[j]);
                }
        __syncthreads();
    }
    *ns_out = ns;
    *acc_out = acc;
}

// FindCoreDensity of the cores cl[0 .. ntask), one workgroup each (slot =
// block).  Members go to the member arena (offset memoff[c]); a core whose
// sweep outgrows the queue or the arena is left as kGiant.
__global__ __launch_bounds__(kBlkT) void core_block_k(CoreArgs a, Pool pool, const int *cl,
                                                      int ntask, int *arena, long long acap,
                                                      unsigned long long *aused,
                                                      long long *memoff) {
    __shared__ BlockLds L;
    const int i = blockIdx.x;
    if (i >= ntask) return;
    const int c = cl[i];
    unsigned *vb = pool.bits + (long long)i * pool.nwords;
    int *q = pool.q + (long long)i * pool.cap;
    const int root = a.peak[c];
    const float gmin = ord_val(*a.kmin) < 1.e23f ? ord_val(*a.kmin) : 1.e23f;
    const float minden = (float)(kDenFloor > (double)gmin ? kDenFloor : (double)gmin);
    float up = a.den[root], dn = minden, t;
    unsigned char st = kOk;
    int steps = 0;
    do {
        t = core_trial(up, dn);
        const long long r = block_sweep(L, kTrial, root, t, a.den, a.nbr, a.k, a.pk, a.stale,
                                        vb, q, pool.cap, nullptr);
        if (r == kSweepBig) {
            st = kGiant;
            break;
        }
        if (r == kSweepHit) dn = t;
        else up = t;
        if (++steps > kMaxSearch) {
            st = kSearchFail;
            break;
        }
    } while (search_more(up, dn, t));
    long long n = 0;
    bool fhit = false;
    if (st == kOk) {
        n = block_sweep(L, kFinal, root, up, a.den, a.nbr, a.k, a.pk, a.stale, vb, q, pool.cap,
                        &fhit);
        if (n == kSweepBig) st = kGiant;
        else if (fhit) st = kQuirk;
    }
    __shared__ long long off;
    if (st == kOk || st == kQuirk) {
        if (threadIdx.x == 0) {
            const unsigned long long o = atomicAdd(aused, (unsigned long long)n);
            off = (long long)o + n <= acap ? (long long)o : -1;
        }
        __syncthreads();
        if (off < 0) st = kGiant;
    }
    if (st != kOk && st != kQuirk) {
        if (threadIdx.x == 0) a.status[c] = st;
        return;
    }
    for (long long j = threadIdx.x; j < n; j += kBlkT) arena[off + j] = q[j];
    int ns = 0;
    float acc = 0.0f;
    if (st == kOk) block_star_sum(L, q, n, a.type, a.mass, &ns, &acc);
    if (threadIdx.x == 0) {
        memoff[c] = off;
        a.status[c] = st == kOk ? kMid : kQuirk;
        a.count[c] = (int)n;
        a.coredensity[c] = up;
        a.nummem[c] = st == kOk ? (int)n : 0;
        a.numstar[c] = ns;
        a.starmass[c] = acc;
    }
}

// Members of the final cores of [c0, c1): from the wave lists (kOk) or the
// arena (kMid).
__global__ void core_apply2_k(const unsigned char *status, const int *count, const int *memb,
                              const int *arena, const long long *memoff, int c0, int c1,
                              unsigned char *flag, int *haloid) {
    const int c = c0 + (int)blockIdx.x;
    if (c >= c1) return;
    const unsigned char st = status[c];
    if (st != kOk && st != kMid) return;
    const int n = count[c];
    const int *m = st == kOk ? memb + (long long)c * kWaveQ : arena + memoff[c];
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int v = m[j];
        flag[v] = (unsigned char)((flag[v] | kWpCore | kWpBound) & ~kWpRemaining);
        atomicMax(&haloid[v], c);
    }
}

// Shell sweeps of the batch entries ent[0 .. ntask) (roots peak[batch[e]]),
// one workgroup each; their lists stay in the pool slots.
__global__ __launch_bounds__(kBlkT) void shell_block_k(const float *den, const long long *nbr,
                                                       int k, const unsigned char *pk,
                                                       const int *peak, const int *batch,
                                                       const int *ent, int ntask, float t,
                                                       Pool pool, unsigned char *status,
                                                       int *count, int *slot_of) {
    __shared__ BlockLds L;
    const int i = blockIdx.x;
    if (i >= ntask) return;
    const int e = ent[i];
    unsigned *vb = pool.bits + (long long)i * pool.nwords;
    int *q = pool.q + (long long)i * pool.cap;
    const long long n = block_sweep(L, kShell, peak[batch[e]], t, den, nbr, k, pk, nullptr, vb,
                                    q, pool.cap, nullptr);
    if (threadIdx.x == 0) {
        if (n == kSweepBig) {
            status[e] = kGiant;
        } else {
            status[e] = kMid;
            count[e] = (int)n;
            slot_of[e] = i;
        }
    }
}

// ---- device-wide shell sweeps of several roots at once -------------------
// Up to kMaxRoots shell sweeps advance level by level together: the level's
// frontier nodes of all roots form one index space (root r's nodes at
// off[r] .. off[r+1]); a candidate claims its particle in its root's claim
// array with its (node, slot) rank; each node counts its winners, an
// exclusive scan of the counts places them, and the winners are appended to
// their root's queue in (node, slot) order -- each root's FIFO order.  The
// host queues levels in groups without waiting; every kernel reads the
// level's size from device memory and exits when all sweeps are done.

constexpr int kMaxRoots = 16;
constexpr int kMlBlocksPerRoot = 256;  // grid x of the per-root level kernels
constexpr int kMlScanBlocks = 1024;
constexpr int kMlLevelsPerSync = 16;
constexpr long long kMlCandCap = 1LL << 27;  // candidates of a level kept between passes

struct MultiSweep {
    int M;
    long long np, nwords;
    int *bq;          // [M][np]
    unsigned *vb;     // [M][nwords]
    int *claim;       // [M][np], INT_MAX idle
    long long *lo, *hi, *off;  // [M], [M], [M + 1]
    unsigned *cnt, *base;      // [M * np] per frontier node
    unsigned *bsum;            // [kMlScanBlocks]
    int *cv;                   // [kMlCandCap]: a level's candidate particles (-1 none)
    int *keep;                 // 1 when this level's candidates fit in cv
    int *done;
    const int *roots;          // [M] root particles
    unsigned *above;           // [nwords]: bit v set when den[v] > the sweep's threshold
    float above_t;             // the threshold `above` holds (host copy), NaN: none
};

__global__ void ml_above_k(const float *den, long long np, float t, unsigned *above) {
    const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w * 32 >= np) return;
    unsigned bits = 0u;
    const long long v0 = w * 32;
    const int nv = np - v0 < 32 ? (int)(np - v0) : 32;
    for (int b = 0; b < nv; b++) bits |= (den[v0 + b] > t ? 1u : 0u) << b;
    above[w] = bits;
}

// The next level's plan (thread 0): each root's frontier [lo, hi) at
// off[r] of the level's node space, whether its candidates fit in cv, and
// whether every sweep is done.
__device__ void ml_plan(MultiSweep &a) {
    long long o = 0;
    for (int r = 0; r < a.M; r++) {
        a.off[r] = o;
        o += a.hi[r] - a.lo[r];
    }
    a.off[a.M] = o;
    *a.keep = o * 32 <= kMlCandCap ? 1 : 0;
    if (o == 0) *a.done = 1;
}

// The roots (one lane each), then the first level's plan.
__global__ void ml_init_k(MultiSweep a) {
    const int r = threadIdx.x;
    if (r < a.M) {
        const int root = a.roots[r];
        a.bq[(long long)r * a.np] = root;
        atomicOr(&a.vb[(long long)r * a.nwords + (root >> 5)], 1u << (root & 31));
        a.lo[r] = 0;
        a.hi[r] = 1;
    }
    if (r == 0) *a.done = 0;
    __syncthreads();
    if (r == 0) ml_plan(a);
}

// Candidate j = node * 32 + slot of root r's level: its particle, or -1.
__device__ __forceinline__ int ml_cand(const MultiSweep &a, int r, long long j, float t,
                                       const float *den, const long long *nbr, int k) {
    const int s = (int)(j & 31);
    if (s >= k) return -1;
    const int u = a.bq[(long long)r * a.np + a.lo[r] + (j >> 5)];
    const int v = (int)nbr[(long long)u * k + s];
    if (!((a.above[v >> 5] >> (v & 31)) & 1u)) return -1;  // den[v] > t
    if ((a.vb[(long long)r * a.nwords + (v >> 5)] >> (v & 31)) & 1u) return -1;
    return v;
}

// Root r = blockIdx.y; its candidates j in [0, nf_r * 32), global index
// off[r] * 32 + j.  The claim pass keeps each candidate's particle in cv
// when the level fits, so the two later passes read it back coalesced.
__global__ void ml_claim_k(MultiSweep a, float t, const float *den, const long long *nbr, int k) {
    if (*a.done) return;
    const int r = blockIdx.y;
    if (r >= a.M) return;
    const long long n = (a.hi[r] - a.lo[r]) * 32, g0 = a.off[r] * 32;
    const bool keep = *a.keep != 0;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (long long)gridDim.x * blockDim.x) {
        const int v = ml_cand(a, r, j, t, den, nbr, k);
        if (keep) a.cv[g0 + j] = v;
        if (v >= 0) atomicMin(&a.claim[(long long)r * a.np + v], (int)j);
    }
}

// Winners of each node: count (count pass) or append (push pass).  The
// stride is a multiple of 64, so a wave always holds two whole nodes.
template <bool kPush>
__global__ void ml_win_k(MultiSweep a, float t, const float *den, const long long *nbr, int k) {
    if (*a.done) return;
    const int r = blockIdx.y;
    if (r >= a.M) return;
    const long long n = (a.hi[r] - a.lo[r]) * 32, g0 = a.off[r] * 32;
    const bool keep = *a.keep != 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    int *claim = a.claim + (long long)r * a.np;
    for (long long j0 = (long long)blockIdx.x * blockDim.x; j0 < n; j0 += stride) {
        const long long j = j0 + threadIdx.x;
        int v = -1;
        bool win = false;
        if (j < n) {
            v = keep ? a.cv[g0 + j] : ml_cand(a, r, j, t, den, nbr, k);
            win = v >= 0 && claim[v] == (int)j;
        }
        const unsigned long long m = __ballot(win);
        const int lane = threadIdx.x & 63;
        const unsigned half = (unsigned)(m >> (lane & 32));
        if (!kPush) {
            if (j < n && (lane & 31) == 0) a.cnt[a.off[r] + (j >> 5)] = __popc(half);
        } else if (win) {
            const long long pos = a.base[a.off[r] + (j >> 5)] - a.base[a.off[r]] +
                                  __popc(half & ((1u << (lane & 31)) - 1u));
            a.bq[(long long)r * a.np + a.hi[r] + pos] = v;
            atomicOr(&a.vb[(long long)r * a.nwords + (v >> 5)], 1u << (v & 31));
            claim[v] = INT_MAX;
        }
    }
}

// Exclusive scan of cnt[0 .. off[M]) into base, sized on the device.
__global__ __launch_bounds__(256) void ml_scan1_k(MultiSweep a) {
    if (*a.done) return;
    const long long n = a.off[a.M];
    const long long seg = (n + kMlScanBlocks - 1) / kMlScanBlocks;
    const long long b0 = blockIdx.x * seg, b1 = b0 + seg < n ? b0 + seg : n;
    unsigned acc = 0;
    for (long long j = b0 + threadIdx.x; j < b1; j += 256) acc += a.cnt[j];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
    __shared__ unsigned w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) a.bsum[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}

__global__ __launch_bounds__(1024) void ml_scan2_k(MultiSweep a) {
    if (*a.done) return;
    __shared__ unsigned sh[1024];
    const unsigned v = a.bsum[threadIdx.x];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const unsigned add = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    a.bsum[threadIdx.x] = sh[threadIdx.x] - v;
}

__global__ __launch_bounds__(256) void ml_scan3_k(MultiSweep a) {
    if (*a.done) return;
    const long long n = a.off[a.M];
    const long long seg = (n + kMlScanBlocks - 1) / kMlScanBlocks;
    const long long b0 = blockIdx.x * seg, b1 = b0 + seg < n ? b0 + seg : n;
    __shared__ unsigned w[4];
    unsigned carry = a.bsum[blockIdx.x];
    for (long long c0 = b0; c0 < b1; c0 += 256) {
        const long long j = c0 + threadIdx.x;
        const unsigned x = j < b1 ? a.cnt[j] : 0u;
        // inclusive scan within the wave, then across the 4 waves
        unsigned inc = x;
        const int lane = threadIdx.x & 63;
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned y = __shfl_up(inc, o);
            if (lane >= o) inc += y;
        }
        if (lane == 63) w[threadIdx.x >> 6] = inc;
        __syncthreads();
        unsigned before = 0, tot = 0;
        for (int q = 0; q < 4; q++) {
            if (q < (int)(threadIdx.x >> 6)) before += w[q];
            tot += w[q];
        }
        if (j < b1) a.base[j] = carry + before + inc - x;
        carry += tot;
        __syncthreads();
    }
}

// The level's appended winners close each root's frontier (one lane per
// root), then the next level's plan.
__global__ void ml_advance_k(MultiSweep a) {
    if (*a.done) return;
    const int r = threadIdx.x;
    if (r < a.M) {
        const long long f0 = a.off[r], f1 = a.off[r + 1];
        const long long c =
            f1 > f0 ? (long long)a.base[f1 - 1] + a.cnt[f1 - 1] - a.base[f0] : 0;
        a.lo[r] = a.hi[r];
        a.hi[r] += c;
    }
    __syncthreads();
    if (r == 0) ml_plan(a);
}

hipError_t multi_bufs(MultiSweep &a, long long np, Workspace &ws, hipStream_t s) {
    a.np = np;
    a.nwords = (np + 31) / 32;
    a.bq = ws.alloc_n<int>((size_t)kMaxRoots * np);
    a.vb = ws.alloc_n<unsigned>((size_t)kMaxRoots * a.nwords);
    a.claim = ws.alloc_n<int>((size_t)kMaxRoots * np);
    a.lo = ws.alloc_n<long long>(kMaxRoots);
    a.hi = ws.alloc_n<long long>(kMaxRoots);
    a.off = ws.alloc_n<long long>(kMaxRoots + 1);
    a.cnt = ws.alloc_n<unsigned>((size_t)kMaxRoots * np);
    a.base = ws.alloc_n<unsigned>((size_t)kMaxRoots * np);
    a.bsum = ws.alloc_n<unsigned>(kMlScanBlocks);
    a.cv = ws.alloc_n<int>(kMlCandCap);
    a.keep = ws.alloc_n<int>(1);
    a.done = ws.alloc_n<int>(1);
    int *roots = ws.alloc_n<int>(kMaxRoots);
    a.roots = roots;
    a.above = ws.alloc_n<unsigned>((size_t)a.nwords);
    a.above_t = __builtin_nanf("");
    if (!a.bq || !a.vb || !a.claim || !a.lo || !a.hi || !a.off || !a.cnt || !a.base || !a.bsum ||
        !a.cv || !a.keep || !a.done || !roots || !a.above)
        return hipErrorOutOfMemory;
    PGALF_HIP_TRY(fill_u32(a.vb, kMaxRoots * a.nwords, 0u, s));
    return fill_u32(reinterpret_cast<unsigned *>(a.claim), kMaxRoots * np, 0x7fffffffu, s);
}

// Shell sweeps (threshold t) of the M <= kMaxRoots roots; on return the
// queues hold n[r] particles each (host array) and the visited bits are
// clear again.
hipError_t multi_sweep(MultiSweep &a, const int *roots_h, int M, float t, const float *den,
                       const long long *nbr, int k, long long *n_h, hipStream_t s) {
    a.M = M;
    PGALF_HIP_TRY(hipMemcpyAsync(const_cast<int *>(a.roots), roots_h, sizeof(int) * M,
                                 hipMemcpyHostToDevice, s));
    if (!(a.above_t == t)) {  // the threshold's bitmap (a 512 KB bit set at 4M
                              // particles, where den is 16 MB), once per threshold
        ml_above_k<<<nblk(a.nwords, 256), 256, 0, s>>>(den, a.np, t, a.above);
        a.above_t = t;
    }
    ml_init_k<<<1, 64, 0, s>>>(a);
    for (;;) {
        for (int l = 0; l < kMlLevelsPerSync; l++) {
            const dim3 grid(kMlBlocksPerRoot, M);
            ml_claim_k<<<grid, 256, 0, s>>>(a, t, den, nbr, k);
            ml_win_k<false><<<grid, 256, 0, s>>>(a, t, den, nbr, k);
            ml_scan1_k<<<kMlScanBlocks, 256, 0, s>>>(a);
            ml_scan2_k<<<1, 1024, 0, s>>>(a);
            ml_scan3_k<<<kMlScanBlocks, 256, 0, s>>>(a);
            ml_win_k<true><<<grid, 256, 0, s>>>(a, t, den, nbr, k);
            ml_advance_k<<<1, 64, 0, s>>>(a);
        }
        int done = 0;
        PGALF_HIP_TRY(to_host(&done, a.done, 1, s));
        if (done) break;
    }
    PGALF_HIP_TRY(to_host(n_h, a.hi, M, s));
    for (int r = 0; r < M; r++)
        bits_clear_k<<<nblk(n_h[r], 256), 256, 0, s>>>(a.bq + (long long)r * a.np, n_h[r],
                                                       a.vb + (long long)r * a.nwords);
    return hipGetLastError();
}

// ---- shells ---------------------------------------------------------------

struct ShellState {
    long long nshell, npart, ncores;
    int stop;      // batch position where the resolve walk stopped
    int err;       // 1: shells > maxshell, 2: core never enclosed, 3: cores > maxcores
    int lastroot;  // core of the last root applied
    int pad;
};

struct ShellArgs {
    const unsigned char *pk;
    const int *core_of;
    unsigned char *flag;
    int *haloid;
    unsigned char *enclosed;  // [numcore], per threshold
    long long *part_ptr;
    int *part;
    long long *core_ptr;
    int *cores;
    long long maxshell, maxcores;
    ShellState *st;
};

__global__ void shell_cand_k(const float *coredensity, const unsigned char *enclosed, int c0,
                             int numcore, float t, unsigned *flags) {
    const int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c < numcore) flags[c - c0] = (coredensity[c] > t && !enclosed[c]) ? 1u : 0u;
}

__global__ void shell_batch_k(const unsigned *flags, const unsigned *pos, int c0, int numcore,
                              int cap, int *batch, int *lastidx) {
    const int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numcore || !flags[c - c0]) return;
    const unsigned p = pos[c - c0];
    if (p < (unsigned)cap) batch[p] = c;
    if (p == (unsigned)cap - 1) *lastidx = c;
}

__global__ __launch_bounds__(64) void shell_wave_k(const float *den, const long long *nbr, int k,
                                                   const unsigned char *pk, const int *peak,
                                                   const int *batch, int nb, float t,
                                                   unsigned char *status, int *memb,
                                                   int *count) {
    __shared__ WaveLds L;
    const int i = (int)blockIdx.x;
    if (i >= nb) return;
    const int root = peak[batch[i]];
    const int n = wave_sweep(L, kShell, root, t, den, nbr, k, pk, nullptr, nullptr);
    if (n == kSweepBig) {
        if (threadIdx.x == 0) status[i] = kBig;
        return;
    }
    int *out = memb + (long long)i * kWaveQ;
    for (int j = threadIdx.x; j < n; j += 64) out[j] = L.q[j];
    if (threadIdx.x == 0) {
        status[i] = kOk;
        count[i] = n;
    }
}

// Block-wide exclusive prefix of a flag (256 threads); *total gets the sum.
__device__ __forceinline__ int block_prefix(bool f, int *wtot, int &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long m = __ballot(f);
    const int in = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wtot[w] = __popcll(m);
    __syncthreads();
    int off = 0;
    total = 0;
    for (int j = 0; j < 4; j++) {
        if (j < w) off += wtot[j];
        total += wtot[j];
    }
    __syncthreads();
    return off + in;
}

// Apply one root's sweep list (in sweep order) with one 256-thread block:
// its peaks' cores enclosed and listed, its free particles made the shell.
// Returns false (every thread) on an error, recorded in the state.
__device__ bool shell_apply_block(const ShellArgs &a, int c, const int *list, int n, int *wtot,
                                  int *serr) {
    ShellState *st = a.st;
    long long nc = st->ncores, npart = st->npart;
    __syncthreads();
    for (int b = 0; b < n; b += 256) {
        const int j = b + (int)threadIdx.x;
        const int v = j < n ? list[j] : -1;
        const bool isp = v >= 0 && a.pk[v] && a.core_of[v] != INT_MAX;
        const bool isf = v >= 0 && !(a.flag[v] & (kWpShell | kWpCore));
        int tp, tf;
        const int pp = block_prefix(isp, wtot, tp);
        const int pf = block_prefix(isf, wtot, tf);
        if (nc + tp > a.maxcores) {
            if (threadIdx.x == 0) st->err = 3;
            return false;
        }
        if (isp) {
            const int cc = a.core_of[v];
            a.cores[nc + pp] = cc;
            a.enclosed[cc] = 1;
        }
        if (isf) {
            a.part[npart + pf] = v;
            a.flag[v] = (unsigned char)((a.flag[v] | kWpShell) & ~kWpRemaining);
            a.haloid[v] = kNotMember;
        }
        nc += tp;
        npart += tf;
        __syncthreads();
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *serr = 0;
        if (!a.enclosed[c]) {
            *serr = st->err = 2;
        } else if (st->nshell >= a.maxshell) {
            *serr = st->err = 1;
        } else {
            st->nshell++;
            st->ncores = nc;
            st->npart = npart;
            a.part_ptr[st->nshell] = npart;
            a.core_ptr[st->nshell] = nc;
            st->lastroot = c;
        }
    }
    __syncthreads();
    return *serr == 0;
}

__global__ __launch_bounds__(256) void shell_resolve_k(ShellArgs a, const int *batch, int nb,
                                                       int start, const unsigned char *status,
                                                       const int *memb, const int *count,
                                                       const int *slot_of, Pool pool) {
    __shared__ int wtot[4];
    __shared__ int serr;
    for (int i = start; i < nb; i++) {
        const int c = batch[i];
        if (a.enclosed[c]) continue;
        const unsigned char st = status[i];
        if (st != kOk && st != kMid) {
            if (threadIdx.x == 0) a.st->stop = i;
            return;
        }
        const int *list = st == kOk ? memb + (long long)i * kWaveQ
                                    : pool.q + (long long)slot_of[i] * pool.cap;
        if (!shell_apply_block(a, c, list, count[i], wtot, &serr)) return;
    }
    if (threadIdx.x == 0) a.st->stop = nb;
}

// Device-wide apply of a long root list: flags, two scans, then the writes.
__global__ void big_flags_k(const int *list, long long n, const unsigned char *pk,
                            const int *core_of, const unsigned char *flag, unsigned *fp,
                            unsigned *ff) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int v = list[j];
    fp[j] = (pk[v] && core_of[v] != INT_MAX) ? 1u : 0u;
    ff[j] = (flag[v] & (kWpShell | kWpCore)) ? 0u : 1u;
}

__global__ void big_write_k(ShellArgs a, const int *list, long long n, const unsigned *fp,
                            const unsigned *pp, const unsigned *ff, const unsigned *pf,
                            const unsigned *tot) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || a.st->err) return;
    if (a.st->ncores + tot[0] > a.maxcores) return;
    const int v = list[j];
    if (fp[j]) {
        const int cc = a.core_of[v];
        a.cores[a.st->ncores + pp[j]] = cc;
        a.enclosed[cc] = 1;
    }
    if (ff[j]) {
        a.part[a.st->npart + pf[j]] = v;
        a.flag[v] = (unsigned char)((a.flag[v] | kWpShell) & ~kWpRemaining);
        a.haloid[v] = kNotMember;
    }
}

__global__ void big_fin_k(ShellArgs a, int c, const unsigned *tot) {
    ShellState *st = a.st;
    if (st->err) return;
    if (st->ncores + tot[0] > a.maxcores) {
        st->err = 3;
    } else if (!a.enclosed[c]) {
        st->err = 2;
    } else if (st->nshell >= a.maxshell) {
        st->err = 1;
    } else {
        st->ncores += tot[0];
        st->npart += tot[1];
        st->nshell++;
        a.part_ptr[st->nshell] = st->npart;
        a.core_ptr[st->nshell] = st->ncores;
        st->lastroot = c;
    }
}

// SaveRemainingParticles2LastShell: particles in no core and no shell.
__global__ void rest_flags_k(const unsigned char *flag, long long np, unsigned *fr) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) fr[i] = (flag[i] & (kWpCore | kWpShell)) ? 0u : 1u;
}

__global__ void rest_write_k(ShellArgs a, long long np, const unsigned *fr, const unsigned *pr) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np || !fr[i] || a.st->err) return;
    a.part[a.st->npart + pr[i]] = (int)i;
    a.flag[i] = (unsigned char)((a.flag[i] | kWpShell) & ~kWpRemaining);
    a.haloid[i] = kNotMember;
}

__global__ void rest_cores_k(ShellArgs a, int numcore, const unsigned *tot) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numcore || tot[0] == 0 || a.st->nshell != 0 || a.st->err) return;
    if (a.st->ncores + numcore > a.maxcores || a.maxshell < 1) return;
    a.cores[a.st->ncores + c] = c;
}

__global__ void rest_fin_k(ShellArgs a, int numcore, const unsigned *tot) {
    ShellState *st = a.st;
    if (st->err || tot[0] == 0) return;
    if (st->nshell == 0) {
        if (a.maxshell < 1) {
            st->err = 1;
            return;
        }
        if (st->ncores + numcore > a.maxcores) {
            st->err = 3;
            return;
        }
        st->ncores += numcore;
        st->nshell = 1;
        a.core_ptr[1] = st->ncores;
    }
    st->npart += tot[0];
    a.part_ptr[st->nshell] = st->npart;
}

__global__ void visited_mark_k(const int *list, long long n, unsigned char *flag) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) flag[list[j]] |= kWpVisited;
}

__global__ void lastroot_reset_k(ShellState *st) { st->lastroot = -1; }

__global__ void set_status_k(unsigned char *status, const int *ent, int n, unsigned char v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) status[ent[i]] = v;
}

__global__ void shell_state_init_k(ShellState *st, long long *part_ptr, long long *core_ptr) {
    st->nshell = st->npart = st->ncores = 0;
    st->stop = 0;
    st->err = 0;
    st->lastroot = -1;
    part_ptr[0] = 0;
    core_ptr[0] = 0;
}

__global__ void fill_u8_k(unsigned char *p, long long n, unsigned char v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}



// ---- catalogue watershed --------------------------------------------------
// A catalogue's halos up to a size limit: one 1024-thread workgroup per halo
// (halos taken from a counter, largest first) runs FindCoreDensity, then the
// shells, in the reference's own order -- cores one after another, roots one
// after another -- with the workgroup sweeps; so the dropped cores' visited
// marks and the last-core member ids come out as one thread leaves them.
// Every index of a halo is local to it: its particles [off[h], off[h+1]),
// its cores [coff[h], coff[h+1]) with local peak indices, its lists local
// (as the batched stage writes them).

constexpr int kMaxShellDivide = 15;

struct CatArgs {
    const long long *off, *coff;  // [nhalo + 1]
    const int *order;             // the halos run here, largest first
    int ntask;
    int *next;                    // work counter
    const float *den;
    const long long *nbr;
    int k;
    const int *type;
    const double *mass;
    const int *peak;              // per core: local peak particle
    float *coredensity;
    int *nummem, *numstar;
    float *starmass;
    int *haloid;                  // per particle: local core or -9999
    unsigned char *flag;
    unsigned char *pk;            // per particle: 1 for a peak
    int *core_of;                 // per particle: first local core of that peak
    unsigned *bits;               // [slots][2 * nw]: visited bits, stale bits
    int *q;                       // [slots][qcap]
    long long nw, qcap;
    int *herr;                    // per halo: 0 done, else left to the one-halo path
    const float *thr;             // [nhalo][15]
    const int *nthr;              // [nhalo]
    unsigned char *enclosed;      // per core
    long long *part_ptr, *core_ptr;  // halo h at sbase[h]
    const long long *sbase;
    int *part;                    // halo h at off[h]
    int *cores;                   // halo h at cbase[h], ccap[h] entries
    const long long *cbase, *ccap;
    long long *nshell;            // [nhalo]
};

enum : int { kCatBig = 1, kCatSearch = 2, kCatBadPeak = 3, kCatCores = 4, kCatEnclose = 5 };

__device__ __forceinline__ int cat_task(int *next, int *s_task) {
    if (threadIdx.x == 0) *s_task = atomicAdd(next, 1);
    __syncthreads();
    const int t = *s_task;
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kBlkT) void cat_fcd_k(CatArgs a) {
    __shared__ BlockLds L;
    __shared__ int s_task, s_bad;
    __shared__ unsigned s_min;
    unsigned *vb = a.bits + (long long)blockIdx.x * 2 * a.nw;
    unsigned *stale = vb + a.nw;
    int *q = a.q + (long long)blockIdx.x * a.qcap;
    const int tid = threadIdx.x;
    for (;;) {
        const int ti = cat_task(a.next, &s_task);
        if (ti >= a.ntask) break;
        const int h = a.order[ti];
        const long long p0 = a.off[h], c0 = a.coff[h];
        const int n = (int)(a.off[h + 1] - p0), nc = (int)(a.coff[h + 1] - c0);
        const float *den = a.den + p0;
        const long long *nbr = a.nbr + p0 * a.k;
        unsigned char *flag = a.flag + p0, *pk = a.pk + p0;
        int *haloid = a.haloid + p0;
        if (tid == 0) {
            s_bad = 0;
            s_min = 0xffffffffu;
        }
        for (int i = tid; i < n; i += kBlkT) {
            flag[i] = 0;
            haloid[i] = kNotMember;
            pk[i] = 0;
        }
        for (long long w = tid; w < (n + 31) / 32; w += kBlkT) stale[w] = 0u;
        __syncthreads();
        unsigned lo = 0xffffffffu;
        for (int i = tid; i < n; i += kBlkT) {
            const unsigned kk = ord_key(den[i]);
            lo = kk < lo ? kk : lo;
        }
        atomicMin(&s_min, lo);
        for (int c = tid; c < nc; c += kBlkT) {
            const int p = a.peak[c0 + c];
            if (p < 0 || p >= n) {
                s_bad = 1;
            } else {
                pk[p] = 1;
                flag[p] = kWpPeak;
            }
        }
        __syncthreads();
        if (s_bad) {
            if (tid == 0) a.herr[h] = kCatBadPeak;
            continue;
        }
        float gmin = ord_val(s_min);
        gmin = gmin < 1.e23f ? gmin : 1.e23f;
        const float minden = (float)(kDenFloor > (double)gmin ? kDenFloor : (double)gmin);
        int err = 0;
        for (int c = 0; c < nc && !err; c++) {
            const int root = a.peak[c0 + c];
            float up = den[root], dn = minden, t;
            int steps = 0;
            do {
                t = core_trial(up, dn);
                const long long r = block_sweep(L, kTrial, root, t, den, nbr, a.k, pk, stale, vb,
                                                q, a.qcap, nullptr);
                if (r == kSweepBig) {
                    err = kCatBig;
                    break;
                }
                if (r == kSweepHit) dn = t;
                else up = t;
                if (++steps > kMaxSearch) {
                    err = kCatSearch;
                    break;
                }
            } while (search_more(up, dn, t));
            if (err) break;
            bool fhit = false;
            const long long nf = block_sweep(L, kFinal, root, up, den, nbr, a.k, pk, stale, vb, q,
                                             a.qcap, &fhit);
            if (nf == kSweepBig) {
                err = kCatBig;
                break;
            }
            int ns = 0;
            float acc = 0.0f;
            if (fhit) {
                // the dropped core's sweep stays visited for the later cores
                for (long long j = tid; j < nf; j += kBlkT) {
                    const int v = q[j];
                    atomicOr(&stale[v >> 5], 1u << (v & 31));
                }
            } else {
                for (long long j = tid; j < nf; j += kBlkT) {
                    const int v = q[j];
                    flag[v] = (unsigned char)((flag[v] | kWpCore | kWpBound) & ~kWpRemaining);
                    haloid[v] = c;
                }
                block_star_sum(L, q, nf, a.type + p0, a.mass + p0, &ns, &acc);
            }
            if (tid == 0) {
                a.coredensity[c0 + c] = up;
                a.nummem[c0 + c] = fhit ? 0 : (int)nf;
                a.numstar[c0 + c] = ns;
                a.starmass[c0 + c] = acc;
            }
            __syncthreads();
        }
        if (tid == 0) a.herr[h] = err;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlkT) void cat_shells_k(CatArgs a) {
    __shared__ BlockLds L;
    __shared__ int s_task, s_root, s_err;
    __shared__ long long s_nshell, s_npart, s_ncores;
    unsigned *vb = a.bits + (long long)blockIdx.x * 2 * a.nw;
    int *q = a.q + (long long)blockIdx.x * a.qcap;
    const int tid = threadIdx.x;
    for (;;) {
        const int ti = cat_task(a.next, &s_task);
        if (ti >= a.ntask) break;
        const int h = a.order[ti];
        if (a.herr[h]) continue;  // uniform: every thread reads the same word
        const long long p0 = a.off[h], c0 = a.coff[h];
        const int n = (int)(a.off[h + 1] - p0), nc = (int)(a.coff[h + 1] - c0);
        const float *den = a.den + p0;
        const long long *nbr = a.nbr + p0 * a.k;
        unsigned char *flag = a.flag + p0, *pk = a.pk + p0, *enc = a.enclosed + c0;
        int *haloid = a.haloid + p0, *core_of = a.core_of + p0, *part = a.part + p0;
        int *cores = a.cores + a.cbase[h];
        long long *pp = a.part_ptr + a.sbase[h], *cp = a.core_ptr + a.sbase[h];
        const float *cd = a.coredensity + c0;
        const long long ccap = a.ccap[h];
        for (int i = tid; i < n; i += kBlkT) core_of[i] = INT_MAX;
        if (tid == 0) {
            s_nshell = s_npart = s_ncores = 0;
            s_err = 0;
            pp[0] = cp[0] = 0;
        }
        __syncthreads();
        for (int c = tid; c < nc; c += kBlkT) atomicMin(&core_of[a.peak[c0 + c]], c);
        __syncthreads();
        const int nsd = a.nthr[h];
        int lastroot = -1;
        float last_t = 0.0f;
        for (int tix = 0; tix < nsd && !s_err; tix++) {
            const float t = a.thr[(long long)h * kMaxShellDivide + tix];
            for (int c = tid; c < nc; c += kBlkT) enc[c] = 0;
            __syncthreads();
            int cursor = 0, thr_root = -1;
            for (;;) {
                // GET_NONENCLOSED_ID (:555-563): the first core above t that
                // no sweep of this threshold reached (the cores before the
                // last root stay enclosed or below t)
                if (tid == 0) s_root = INT_MAX;
                __syncthreads();
                for (int c0b = cursor; c0b < nc; c0b += kBlkT) {
                    const int c = c0b + tid;
                    if (c < nc && cd[c] > t && !enc[c]) atomicMin(&s_root, c);
                    __syncthreads();
                    const int found = s_root;
                    __syncthreads();
                    if (found != INT_MAX) break;
                }
                const int root = s_root;
                __syncthreads();
                if (root == INT_MAX) break;
                cursor = root;
                const long long ns = block_sweep(L, kShell, a.peak[c0 + root], t, den, nbr, a.k,
                                                 pk, nullptr, vb, q, a.qcap, nullptr);
                if (ns == kSweepBig) {
                    if (tid == 0) s_err = kCatBig;
                    __syncthreads();
                    break;
                }
                long long nco = s_ncores, npa = s_npart;
                __syncthreads();
                bool over = false;
                for (long long b = 0; b < ns; b += kBlkT) {
                    const long long j = b + tid;
                    const int v = j < ns ? q[j] : -1;
                    const bool isp = v >= 0 && pk[v] && core_of[v] != INT_MAX;
                    const bool isf = v >= 0 && !(flag[v] & (kWpShell | kWpCore));
                    int tp, tf;
                    const int ppos = block_prefix_1k(isp, L.wtot, tp);
                    const int fpos = block_prefix_1k(isf, L.wtot, tf);
                    if (nco + tp > ccap) {
                        over = true;
                        break;
                    }
                    if (isp) {
                        const int cc = core_of[v];
                        cores[nco + ppos] = cc;
                        enc[cc] = 1;
                    }
                    if (isf) {
                        part[npa + fpos] = v;
                        flag[v] = (unsigned char)((flag[v] | kWpShell) & ~kWpRemaining);
                        haloid[v] = kNotMember;
                    }
                    nco += tp;
                    npa += tf;
                    __syncthreads();
                }
                __syncthreads();
                if (tid == 0) {
                    if (over) {
                        s_err = kCatCores;
                    } else if (!enc[root]) {
                        s_err = kCatEnclose;
                    } else {
                        s_nshell++;
                        s_ncores = nco;
                        s_npart = npa;
                        pp[s_nshell] = npa;
                        cp[s_nshell] = nco;
                    }
                }
                __syncthreads();
                if (s_err) break;
                thr_root = root;
            }
            lastroot = thr_root;
            last_t = t;
        }
        if (!s_err) {
            // SaveRemainingParticles2LastShell (:572-611)
            long long npa = s_npart;
            for (int b = 0; b < n; b += kBlkT) {
                const int i = b + tid;
                const bool isr = i < n && !(flag[i] & (kWpCore | kWpShell));
                int tr;
                const int rpos = block_prefix_1k(isr, L.wtot, tr);
                if (isr) {
                    part[npa + rpos] = i;
                    flag[i] = (unsigned char)((flag[i] | kWpShell) & ~kWpRemaining);
                    haloid[i] = kNotMember;
                }
                npa += tr;
                __syncthreads();
            }
            if (tid == 0 && npa > s_npart) {
                if (s_nshell == 0) {
                    if (s_ncores + nc > ccap) {
                        s_err = kCatCores;
                    } else {
                        for (int c = 0; c < nc; c++) cores[s_ncores + c] = c;
                        s_ncores += nc;
                        s_nshell = 1;
                        cp[1] = s_ncores;
                    }
                }
                s_npart = npa;
                pp[s_nshell] = npa;
            }
            __syncthreads();
        }
        if (!s_err && lastroot >= 0) {
            // the last sweep of the last threshold leaves its marks
            const long long ns = block_sweep(L, kShell, a.peak[c0 + lastroot], last_t, den, nbr,
                                             a.k, pk, nullptr, vb, q, a.qcap, nullptr);
            for (long long j = tid; j < ns; j += kBlkT) flag[q[j]] |= kWpVisited;
        }
        __syncthreads();
        if (tid == 0) {
            a.herr[h] = s_err;
            a.nshell[h] = s_nshell;
        }
        __syncthreads();
    }
}

// Per-halo minimum density and largest core density (ordered keys).
// Minimum density and largest core density of the halos list[0..gridDim.x)
// (ordered keys).
__global__ void cat_minmax_k(const int *list, const long long *off, const long long *coff,
                             const float *den, const float *cd, unsigned *kmin, unsigned *kmax) {
    const int h = list[blockIdx.x];
    unsigned lo = 0xffffffffu, hi = 0u;
    for (long long i = off[h] + threadIdx.x; i < off[h + 1]; i += blockDim.x) {
        const unsigned kk = ord_key(den[i]);
        lo = kk < lo ? kk : lo;
    }
    for (long long c = coff[h] + threadIdx.x; c < coff[h + 1]; c += blockDim.x) {
        const unsigned kk = ord_key(cd[c]);
        hi = kk > hi ? kk : hi;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned a2 = __shfl_down(lo, o), b2 = __shfl_down(hi, o);
        lo = a2 < lo ? a2 : lo;
        hi = b2 > hi ? b2 : hi;
    }
    __shared__ unsigned slo[4], shi[4];
    if ((threadIdx.x & 63) == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            lo = slo[w] < lo ? slo[w] : lo;
            hi = shi[w] > hi ? shi[w] : hi;
        }
        kmin[h] = lo;
        kmax[h] = hi;
    }
}

}  // namespace

// ---------------------------------------------------------------------------

namespace {

// Pool of workgroup sweeps: at most `want` slots, bitmaps + queues within
// about 1 GiB.  A workgroup sweep holds up to 64K particles: past that the
// device-wide sweeps are faster.
hipError_t make_pool(Pool &p, long long np, int want, Workspace &ws, hipStream_t s) {
    p.nwords = (np + 31) / 32;
    p.cap = np + kBlkT < (1LL << 16) ? np + kBlkT : (1LL << 16);
    const long long per = p.nwords * 4 + p.cap * 4;
    long long slots = (1LL << 30) / (per > 0 ? per : 1);
    slots = slots < want ? slots : want;
    p.slots = (int)(slots > 1 ? slots : 1);
    p.bits = ws.alloc_n<unsigned>((size_t)p.slots * p.nwords);
    p.q = ws.alloc_n<int>((size_t)p.slots * p.cap);
    if (!p.bits || !p.q) return hipErrorOutOfMemory;
    return fill_u32(p.bits, (long long)p.slots * p.nwords, 0u, s);
}

float host_val(unsigned k2) {
    const unsigned b = (k2 >> 31) ? (k2 & 0x7fffffffu) : ~k2;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

}  // namespace

hipError_t find_core_density(const float *den, long long np, const long long *nbr, int k,
                             const int *type, const double *mass, const int *peak, int numcore,
                             const CoreOut &o, int *search_fail, Workspace &ws, hipStream_t s) {
    if (search_fail) *search_fail = 0;
    WsScope scope(ws);
    const long long nwords = (np + 31) / 32;
    unsigned char *pk = ws.alloc_n<unsigned char>(np > 0 ? np : 1);
    int *core_of = ws.alloc_n<int>(np > 0 ? np : 1);
    unsigned *stale = ws.alloc_n<unsigned>(nwords > 0 ? nwords : 1);
    unsigned *kk = ws.alloc_n<unsigned>(2);
    const int nc1 = numcore > 0 ? numcore : 1;
    unsigned char *status = ws.alloc_n<unsigned char>(nc1);
    int *memb = ws.alloc_n<int>((size_t)nc1 * kWaveQ);
    int *count = ws.alloc_n<int>(nc1);
    long long *memoff = ws.alloc_n<long long>(nc1);
    int *cl = ws.alloc_n<int>(nc1);
    unsigned long long *aused = ws.alloc_n<unsigned long long>(1);
    if (!pk || !core_of || !stale || !kk || !status || !memb || !count || !memoff || !cl ||
        !aused)
        return hipErrorOutOfMemory;
    if (np <= 0) return hipSuccess;
    ws_init_k<<<nblk(np, 256), 256, 0, s>>>(np, o.flag, o.haloid, core_of);
    if (numcore > 0)
        ws_peaks_k<<<nblk(numcore, 256), 256, 0, s>>>(peak, numcore, np, o.flag, core_of,
                                                      ws.status());
    ws_peakbyte_k<<<nblk(np, 256), 256, 0, s>>>(o.flag, np, pk);
    PGALF_HIP_TRY(fill_u32(stale, nwords, 0u, s));
    ws_minmax_init_k<<<1, 1, 0, s>>>(kk);
    ws_minmax_k<<<nblk(np, 256) < 2048 ? nblk(np, 256) : 2048, 256, 0, s>>>(den, np, kk,
                                                                            nullptr);
    if (numcore <= 0) return hipGetLastError();
    {
        // a peak outside [0, np) would be read by the sweeps: stop here
        unsigned st = 0;
        PGALF_HIP_TRY(to_host(&st, ws.status(), 1, s));
        if (st & DEV_ERR_BAD_INDEX) return hipSuccess;
    }
    CoreArgs a;
    a.den = den;
    a.nbr = nbr;
    a.k = k;
    a.pk = pk;
    a.stale = stale;
    a.type = type;
    a.mass = mass;
    a.peak = peak;
    a.numcore = numcore;
    a.kmin = kk;
    a.coredensity = o.coredensity;
    a.nummem = o.nummem;
    a.numstar = o.numstar;
    a.starmass = o.starmass;
    a.status = status;
    a.memb = memb;
    a.count = count;
    Pool pool{};
    bool have_pool = false;
    LevelBufs lb;
    bool have_lb = false;
    int *arena = nullptr;
    const long long acap = 2 * np + 4096;
    std::vector<unsigned char> hst(numcore);
    std::vector<int> hpeak;
    int start = 0;
    while (start < numcore) {
        if (have_pool) PGALF_HIP_TRY(fill_u64(aused, 1, 0ull, s));  // the arena is free again
        core_wave_k<<<numcore - start, 64, 0, s>>>(a, start);
        PGALF_HIP_TRY(hipGetLastError());
        PGALF_HIP_TRY(to_host(hst.data() + start, status + start, numcore - start, s));
        // the first core dropped by its final sweep ends this pass
        int q = numcore;
        std::vector<int> bigs;
        for (int c = start; c < numcore && q == numcore; c++) {
            if (hst[c] == kQuirk) q = c;
            else if (hst[c] == kBig) bigs.push_back(c);
        }
        // cores too large for the LDS: one workgroup each
        if (!bigs.empty()) {
            if (!have_pool) {
                PGALF_HIP_TRY(make_pool(pool, np, (int)(bigs.size() < 512 ? bigs.size() : 512),
                                        ws, s));
                arena = ws.alloc_n<int>(acap);
                if (!arena) return hipErrorOutOfMemory;
                PGALF_HIP_TRY(fill_u64(aused, 1, 0ull, s));
                have_pool = true;
            }
            for (size_t b = 0; b < bigs.size(); b += pool.slots) {
                const int m = (int)(bigs.size() - b < (size_t)pool.slots ? bigs.size() - b
                                                                          : pool.slots);
                PGALF_HIP_TRY(hipMemcpyAsync(cl, bigs.data() + b, sizeof(int) * m,
                                             hipMemcpyHostToDevice, s));
                core_block_k<<<m, kBlkT, 0, s>>>(a, pool, cl, m, arena, acap, aused, memoff);
                PGALF_HIP_TRY(hipGetLastError());
            }
            PGALF_HIP_TRY(to_host(hst.data() + start, status + start, numcore - start, s));
            for (int c : bigs)
                if (c < q && hst[c] == kQuirk) q = c;
        }
        // giants, in core order, with device-wide sweeps
        for (int c : bigs) {
            if (c > q) break;
            if (hst[c] == kSearchFail) {
                if (search_fail) *search_fail = 1;
                continue;
            }
            if (hst[c] != kGiant) continue;
            if (!have_lb) {
                PGALF_HIP_TRY(level_bufs(lb, np, k, ws, s));
                have_lb = true;
            }
            if (hpeak.empty()) {
                hpeak.resize(numcore);
                PGALF_HIP_TRY(to_host(hpeak.data(), peak, numcore, s));
            }
            unsigned kmin = 0;
            PGALF_HIP_TRY(to_host(&kmin, kk, 1, s));
            float g = host_val(kmin);
            g = g < 1.e23f ? g : 1.e23f;
            const float minden = (float)(kDenFloor > (double)g ? kDenFloor : (double)g);
            const int root = hpeak[c];
            float up = 0.0f, dn = minden, t;
            PGALF_HIP_TRY(to_host(&up, den + root, 1, s));
            int steps = 0;
            bool fail = false;
            do {
                t = (float)(0.5 * (double)(up + dn));
                long long n;
                bool hit;
                PGALF_HIP_TRY(level_sweep(lb, kTrial, root, t, den, nbr, k, pk, stale, &n, &hit,
                                          ws, s));
                if (hit) dn = t;
                else up = t;
                if (++steps > kMaxSearch) {
                    fail = true;
                    break;
                }
            } while (search_more(up, dn, t));
            if (fail) {
                if (search_fail) *search_fail = 1;
                core_scalar_k<<<1, 1, 0, s>>>(c, up, 0, o.nummem, o.numstar, o.starmass,
                                              o.coredensity, status, kSearchFail);
                continue;
            }
            long long n;
            bool hit;
            PGALF_HIP_TRY(level_sweep(lb, kFinal, root, up, den, nbr, k, pk, stale, &n, &hit, ws,
                                      s));
            if (hit) {
                core_scalar_k<<<1, 1, 0, s>>>(c, up, 0, o.nummem, o.numstar, o.starmass,
                                              o.coredensity, status, kQuirk);
                stale_mark_k<<<nblk(n, 256), 256, 0, s>>>(lb.bq, n, stale);
                q = c;
                break;
            }
            core_scalar_k<<<1, 1, 0, s>>>(c, up, -1, o.nummem, o.numstar, o.starmass,
                                          o.coredensity, status, kGiant);
            star_sum_k<<<1, 256, 0, s>>>(lb.bq, n, type, mass, c, o.nummem, o.numstar,
                                         o.starmass);
            list_apply_k<<<nblk(n, 256), 256, 0, s>>>(lb.bq, n, c, o.flag, o.haloid);
        }
        for (int c = start; c < q; c++)
            if (hst[c] == kSearchFail && search_fail) *search_fail = 1;
        // members of the final cores of this pass
        const int c1 = q < numcore ? q : numcore;
        if (c1 > start)
            core_apply2_k<<<c1 - start, 256, 0, s>>>(status, count, memb, arena, memoff, start,
                                                     c1, o.flag, o.haloid);
        if (q < numcore && hst[q] == kQuirk) {
            // the dropped core's sweep stays visited for every later core
            // (a giant's was marked above)
            int n = 0;
            long long off = 0;
            PGALF_HIP_TRY(to_host(&n, count + q, 1, s));
            const bool in_arena = std::find(bigs.begin(), bigs.end(), q) != bigs.end();
            if (in_arena) PGALF_HIP_TRY(to_host(&off, memoff + q, 1, s));
            stale_mark_k<<<nblk(n, 256), 256, 0, s>>>(
                in_arena ? arena + off : memb + (long long)q * kWaveQ, n, stale);
        }
        start = q + 1;
    }
    return hipGetLastError();
}

// nshelldivide of subhalo_den: 10, or min(15, 10 log10(1.5 np)) for np > 1000.
static int shell_divide(long long np) {
    int nsd = 10;
    if (np > 1000) {
        const double a2 = 10.0 * log10((double)np * 1.5);
        nsd = (int)(10.0 > a2 ? 10.0 : a2);
        nsd = nsd < 15 ? nsd : 15;
    }
    return nsd;
}

// The shell thresholds of subhalo_den (subhaloden.mod6.c:2995-3024) from the
// smallest particle density pmin and the largest core density cmax (1e27 /
// -1e27 when there are none), in FP32 variables with glibc's log10 and pow
// as the reference forms them; thr[0 .. return), highest first.
static int thresholds_from(float pmin, float cmax, long long np, float *thr) {
    float minshellden = 1.e27f, maxshellden = -1.e27f;
    maxshellden = maxshellden > cmax ? maxshellden : cmax;
    minshellden = minshellden < pmin ? minshellden : pmin;
    minshellden = (float)(kDenFloor > (double)minshellden ? kDenFloor : (double)minshellden);
    minshellden = (float)log10((double)minshellden + 1.0);
    maxshellden = (float)log10((double)maxshellden + 1.0);
    const int nsd = shell_divide(np);
    for (int i = 0; i < nsd; i++) {
        const float step = (float)(nsd - 1 - i) * (maxshellden - minshellden);
        const float tl = minshellden + step / (float)nsd;
        thr[i] = (float)(pow(10.0, (double)tl) - 1.0);
    }
    return nsd;
}

hipError_t shell_thresholds(const float *den, long long np, const float *coredensity,
                            int numcore, float *thr, int *nthr, Workspace &ws, hipStream_t s) {
    WsScope scope(ws);
    unsigned *kk = ws.alloc_n<unsigned>(4);
    if (!kk) return hipErrorOutOfMemory;
    ws_minmax_init_k<<<1, 1, 0, s>>>(kk);
    ws_minmax_init_k<<<1, 1, 0, s>>>(kk + 2);
    if (np > 0)
        ws_minmax_k<<<nblk(np, 256) < 2048 ? nblk(np, 256) : 2048, 256, 0, s>>>(den, np, kk,
                                                                                nullptr);
    if (numcore > 0)
        ws_minmax_k<<<nblk(numcore, 256), 256, 0, s>>>(coredensity, numcore, nullptr, kk + 3);
    unsigned h[4];
    PGALF_HIP_TRY(to_host(h, kk, 4, s));
    auto val = [](unsigned k2) {
        const unsigned b = (k2 >> 31) ? (k2 & 0x7fffffffu) : ~k2;
        float f;
        memcpy(&f, &b, 4);
        return f;
    };
    const float pmin = np > 0 ? val(h[0]) : 1.e27f;
    const float cmax = numcore > 0 ? val(h[3]) : -1.e27f;
    *nthr = thresholds_from(pmin, cmax, np, thr);
    return hipSuccess;
}


hipError_t core_shells(const float *den, long long np, const long long *nbr, int k,
                       const int *peak, const float *coredensity, int numcore, const float *thr,
                       int nthr, unsigned char *flag, int *haloid, const ShellOut &o,
                       long long *nshell_out, int *err_out, Workspace &ws, hipStream_t s) {
    *nshell_out = 0;
    *err_out = 0;
    WsScope scope(ws);
    const int nc1 = numcore > 0 ? numcore : 1;
    const int B = nc1 < 8192 ? nc1 : 8192;
    const long long np1 = np > 0 ? np : 1;
    unsigned char *pk = ws.alloc_n<unsigned char>(np1);
    int *core_of = ws.alloc_n<int>(np1);
    unsigned char *enclosed = ws.alloc_n<unsigned char>(nc1);
    unsigned *cflags = ws.alloc_n<unsigned>(nc1), *cpos = ws.alloc_n<unsigned>(nc1);
    int *batch = ws.alloc_n<int>(B), *lastidx = ws.alloc_n<int>(1);
    unsigned char *bstatus = ws.alloc_n<unsigned char>(B);
    int *memb = ws.alloc_n<int>((size_t)B * kWaveQ), *count = ws.alloc_n<int>(B);
    int *slot_of = ws.alloc_n<int>(B), *ent = ws.alloc_n<int>(B);
    unsigned *ctot = ws.alloc_n<unsigned>(1);
    ShellState *st = ws.alloc_n<ShellState>(1);
    unsigned *fp = ws.alloc_n<unsigned>(np1), *pp = ws.alloc_n<unsigned>(np1);
    unsigned *ff = ws.alloc_n<unsigned>(np1), *pf = ws.alloc_n<unsigned>(np1);
    unsigned *tot = ws.alloc_n<unsigned>(2);
    if (!pk || !core_of || !enclosed || !cflags || !cpos || !batch || !lastidx || !bstatus ||
        !memb || !count || !slot_of || !ent || !ctot || !st || !fp || !pp || !ff || !pf || !tot)
        return hipErrorOutOfMemory;
    shell_state_init_k<<<1, 1, 0, s>>>(st, o.part_ptr, o.core_ptr);
    if (np <= 0) return hipGetLastError();
    ws_peakbyte_k<<<nblk(np, 256), 256, 0, s>>>(flag, np, pk);
    PGALF_HIP_TRY(fill_u32(reinterpret_cast<unsigned *>(core_of), np, 0x7fffffffu, s));
    if (numcore > 0) {
        core_of_k<<<nblk(numcore, 256), 256, 0, s>>>(peak, numcore, np, core_of, ws.status());
        unsigned sw = 0;
        PGALF_HIP_TRY(to_host(&sw, ws.status(), 1, s));
        if (sw & DEV_ERR_BAD_INDEX) return hipSuccess;
    }
    ShellArgs a;
    a.pk = pk;
    a.core_of = core_of;
    a.flag = flag;
    a.haloid = haloid;
    a.enclosed = enclosed;
    a.part_ptr = o.part_ptr;
    a.part = o.part;
    a.core_ptr = o.core_ptr;
    a.cores = o.cores;
    a.maxshell = o.maxshell;
    a.maxcores = o.maxcores;
    a.st = st;
    LevelBufs lb;
    bool have_lb = false;
    Pool pool{};
    bool have_pool = false;
    std::vector<int> hpeak;
    auto get_hpeak = [&]() -> hipError_t {
        if (!hpeak.empty()) return hipSuccess;
        hpeak.resize(numcore);
        return to_host(hpeak.data(), peak, numcore, s);
    };
    auto get_lb = [&]() -> hipError_t {
        if (have_lb) return hipSuccess;
        have_lb = true;
        return level_bufs(lb, np, k, ws, s);
    };
    std::vector<unsigned char> hst(B), henc(nc1);
    std::vector<int> hbatch(B), hent;
    // once a sweep has outgrown the workgroup tier, the lower thresholds'
    // roots too large for the LDS go straight to the device-wide sweep (the
    // sweeps only grow as the threshold falls)
    bool giant_mode = false;
    MultiSweep ms{};
    bool have_ms = false;
    std::vector<long long> hn(kMaxRoots);
    std::vector<int> hslot(B);
    int last_root = -1;
    float last_thr = 0.0f;
    for (int ti = 0; ti < nthr; ti++) {
        const float t = thr[ti];
        fill_u8_k<<<nblk(nc1, 256), 256, 0, s>>>(enclosed, nc1, 0);
        lastroot_reset_k<<<1, 1, 0, s>>>(st);
        int next = 0;
        int spec = 1;    // workgroup sweeps launched at a time (grows while none is giant)
        int mgroup = 1;  // device-wide sweeps run together (grows after the first)
        while (next < numcore) {
            const int nr = numcore - next;
            shell_cand_k<<<nblk(nr, 256), 256, 0, s>>>(coredensity, enclosed, next, numcore, t,
                                                       cflags);
            {
                WsScope sc(ws);
                PGALF_HIP_TRY(exclusive_scan_u32(cflags, cpos, nr, ctot, ws, s));
            }
            shell_batch_k<<<nblk(nr, 256), 256, 0, s>>>(cflags, cpos, next, numcore, B, batch,
                                                        lastidx);
            unsigned total = 0;
            PGALF_HIP_TRY(to_host(&total, ctot, 1, s));
            if (total == 0) break;
            const int nb = total < (unsigned)B ? (int)total : B;
            if (total > (unsigned)B) {
                int li = 0;
                PGALF_HIP_TRY(to_host(&li, lastidx, 1, s));
                next = li + 1;
            } else {
                next = numcore;
            }
            shell_wave_k<<<nb, 64, 0, s>>>(den, nbr, k, pk, peak, batch, nb, t, bstatus, memb,
                                           count);
            PGALF_HIP_TRY(to_host(hbatch.data(), batch, nb, s));
            int startpos = 0;
            for (;;) {
                shell_resolve_k<<<1, 256, 0, s>>>(a, batch, nb, startpos, bstatus, memb, count,
                                                  slot_of, pool);
                ShellState hs;
                PGALF_HIP_TRY(to_host(&hs, st, 1, s));
                if (hs.err) {
                    *err_out = hs.err;
                    return hipSuccess;
                }
                if (hs.stop >= nb) break;
                const int i = hs.stop;
                PGALF_HIP_TRY(to_host(hst.data() + i, bstatus + i, nb - i, s));
                if (hst[i] == kHuge) {
                    // a root swept device-wide: applied by scans
                    const int c = hbatch[i];
                    const long long n = hn[hslot[i]];
                    const int *list = ms.bq + (long long)hslot[i] * np;
                    big_flags_k<<<nblk(n, 256), 256, 0, s>>>(list, n, pk, core_of, flag, fp, ff);
                    {
                        WsScope sc(ws);
                        PGALF_HIP_TRY(exclusive_scan_u32(fp, pp, n, tot, ws, s));
                        PGALF_HIP_TRY(exclusive_scan_u32(ff, pf, n, tot + 1, ws, s));
                    }
                    big_write_k<<<nblk(n, 256), 256, 0, s>>>(a, list, n, fp, pp, ff, pf, tot);
                    big_fin_k<<<1, 1, 0, s>>>(a, c, tot);
                    startpos = i + 1;
                    continue;
                }
                if (hst[i] == kBig && !giant_mode) {
                    // workgroup sweeps of this root and of the next not yet
                    // enclosed candidates too large for the LDS
                    if (!have_pool) {
                        PGALF_HIP_TRY(make_pool(pool, np, 256, ws, s));
                        have_pool = true;
                    }
                    PGALF_HIP_TRY(to_host(henc.data(), enclosed, nc1, s));
                    const int want = spec < pool.slots ? spec : pool.slots;
                    hent.clear();
                    for (int j = i; j < nb && (int)hent.size() < want; j++)
                        if (hst[j] == kBig && !henc[hbatch[j]]) hent.push_back(j);
                    PGALF_HIP_TRY(hipMemcpyAsync(ent, hent.data(), sizeof(int) * hent.size(),
                                                 hipMemcpyHostToDevice, s));
                    shell_block_k<<<(unsigned)hent.size(), kBlkT, 0, s>>>(
                        den, nbr, k, pk, peak, batch, ent, (int)hent.size(), t, pool, bstatus,
                        count, slot_of);
                    PGALF_HIP_TRY(to_host(hst.data() + i, bstatus + i, nb - i, s));
                    bool giant = false;
                    for (int j : hent) giant |= hst[j] == kGiant;
                    giant_mode |= giant;
                    spec = giant ? 1 : (spec * 4 < pool.slots ? spec * 4 : pool.slots);
                    startpos = i;
                    continue;
                }
                // kGiant, or kBig past the first giant: device-wide sweeps of
                // this root and of the next not yet enclosed such candidates
                giant_mode = true;
                if (!have_ms) {
                    PGALF_HIP_TRY(multi_bufs(ms, np, ws, s));
                    have_ms = true;
                }
                PGALF_HIP_TRY(get_hpeak());
                PGALF_HIP_TRY(to_host(henc.data(), enclosed, nc1, s));
                hent.clear();
                std::vector<int> roots;
                for (int j = i; j < nb && (int)hent.size() < mgroup; j++)
                    if ((hst[j] == kGiant || hst[j] == kBig) && !henc[hbatch[j]]) {
                        hent.push_back(j);
                        roots.push_back(hpeak[hbatch[j]]);
                    }
                PGALF_HIP_TRY(multi_sweep(ms, roots.data(), (int)roots.size(), t, den, nbr, k,
                                          hn.data(), s));
                for (size_t r = 0; r < hent.size(); r++) {
                    hslot[hent[r]] = (int)r;
                    hst[hent[r]] = kHuge;
                }
                PGALF_HIP_TRY(hipMemcpyAsync(ent, hent.data(), sizeof(int) * hent.size(),
                                             hipMemcpyHostToDevice, s));
                set_status_k<<<1, 64, 0, s>>>(bstatus, ent, (int)hent.size(), kHuge);
                mgroup = mgroup * 4 < kMaxRoots ? mgroup * 4 : kMaxRoots;
                startpos = i;
            }
        }
        ShellState hs;
        PGALF_HIP_TRY(to_host(&hs, st, 1, s));
        if (hs.err) {
            *err_out = hs.err;
            return hipSuccess;
        }
        last_root = hs.lastroot;
        last_thr = t;
    }
    // SaveRemainingParticles2LastShell
    rest_flags_k<<<nblk(np, 256), 256, 0, s>>>(flag, np, fp);
    {
        WsScope sc(ws);
        PGALF_HIP_TRY(exclusive_scan_u32(fp, pp, np, tot, ws, s));
    }
    rest_write_k<<<nblk(np, 256), 256, 0, s>>>(a, np, fp, pp);
    if (numcore > 0) rest_cores_k<<<nblk(numcore, 256), 256, 0, s>>>(a, numcore, tot);
    rest_fin_k<<<1, 1, 0, s>>>(a, numcore, tot);
    // the last sweep of the last threshold leaves its particles marked
    // visited (GetShellParticlesPeaks clears the marks before each sweep,
    // not after the last)
    if (last_root >= 0) {
        PGALF_HIP_TRY(get_hpeak());
        PGALF_HIP_TRY(get_lb());
        long long n;
        bool hit;
        PGALF_HIP_TRY(level_sweep(lb, kShell, hpeak[last_root], last_thr, den, nbr, k, pk,
                                  nullptr, &n, &hit, ws, s));
        visited_mark_k<<<nblk(n, 256), 256, 0, s>>>(lb.bq, n, flag);
    }
    ShellState hs;
    PGALF_HIP_TRY(to_host(&hs, st, 1, s));
    *err_out = hs.err;
    *nshell_out = hs.nshell;
    return hipGetLastError();
}


// Helper streams of the catalogue watershed, per device (default
// priority): one for the workgroup halos, and kHelpers more whose threads
// take large halos through the one-halo tiers beside the caller's thread,
// each with a workspace of its own.  An event orders them after the
// caller's stream.
constexpr int kHelpers = 3;
struct SideSet {
    hipStream_t wg = nullptr;
    hipEvent_t ev = nullptr;
    hipStream_t s[kHelpers] = {};
    Workspace *ws[kHelpers] = {};
};

static hipError_t side_set(int dev, SideSet **out) {
    static std::mutex mu;
    static std::map<int, SideSet> sets;
    std::lock_guard<std::mutex> lk(mu);
    SideSet &x = sets[dev];
    if (!x.wg) PGALF_HIP_TRY(hipStreamCreateWithFlags(&x.wg, hipStreamNonBlocking));
    if (!x.ev) PGALF_HIP_TRY(hipEventCreateWithFlags(&x.ev, hipEventDisableTiming));
    for (int i = 0; i < kHelpers; i++) {
        if (!x.s[i]) PGALF_HIP_TRY(hipStreamCreateWithFlags(&x.s[i], hipStreamNonBlocking));
        if (!x.ws[i]) x.ws[i] = new Workspace(dev);
    }
    *out = &x;
    return hipSuccess;
}

__global__ void or_status_k(unsigned *status, unsigned bits) { *status |= bits; }

void catalog_shell_layout(int nhalo, const long long *off, const long long *coff,
                          long long *sbase, long long *cbase, long long *ccap,
                          long long *tot_shell, long long *tot_cores) {
    long long sb = 0, cb = 0;
    for (int h = 0; h < nhalo; h++) {
        const long long n = off[h + 1] - off[h], nc = coff[h + 1] - coff[h];
        const int nsd = shell_divide(n);
        sbase[h] = sb;
        cbase[h] = cb;
        ccap[h] = 16 * nc * (nsd + 1) + 16;  // a core listed by every sweep that reaches it
        sb += nc * nsd + 2;  // maxshell + 1 pointer entries
        cb += ccap[h];
    }
    *tot_shell = sb;
    *tot_cores = cb;
}

hipError_t catalog_watershed(int nhalo, const long long *off_h, const long long *coff_h,
                             const float *den, const long long *nbr, int k, const int *type,
                             const double *mass, const int *peak, const CoreOut &o,
                             const CatShellOut &so, long long small_max, long long *nshell_h,
                             int *herr_h, Workspace &ws, hipStream_t s) {
    WsScope scope(ws);
    std::vector<long long> sbase(nhalo), cbase(nhalo), ccap(nhalo);
    long long tot_s = 0, tot_c = 0;
    catalog_shell_layout(nhalo, off_h, coff_h, sbase.data(), cbase.data(), ccap.data(), &tot_s,
                         &tot_c);
    const long long np = off_h[nhalo], ncore = coff_h[nhalo];
    // the halos for the workgroups (largest first), the rest one by one
    std::vector<int> small, big;
    long long nmax_small = 0;
    for (int h = 0; h < nhalo; h++) {
        const long long n = off_h[h + 1] - off_h[h];
        if (n <= 0) continue;
        if (n <= small_max) {
            small.push_back(h);
            nmax_small = n > nmax_small ? n : nmax_small;
        } else {
            big.push_back(h);
        }
    }
    std::stable_sort(small.begin(), small.end(), [&](int a, int b) {
        return off_h[a + 1] - off_h[a] > off_h[b + 1] - off_h[b];
    });
    long long *doff = ws.alloc_n<long long>(nhalo + 1), *dcoff = ws.alloc_n<long long>(nhalo + 1);
    long long *dsbase = ws.alloc_n<long long>(nhalo > 0 ? nhalo : 1);
    long long *dcbase = ws.alloc_n<long long>(nhalo > 0 ? nhalo : 1);
    long long *dccap = ws.alloc_n<long long>(nhalo > 0 ? nhalo : 1);
    long long *dnshell = ws.alloc_n<long long>(nhalo > 0 ? nhalo : 1);
    int *order = ws.alloc_n<int>(small.size() > 0 ? small.size() : 1);
    int *next = ws.alloc_n<int>(2), *herr = ws.alloc_n<int>(nhalo > 0 ? nhalo : 1);
    unsigned char *pk = ws.alloc_n<unsigned char>(np > 0 ? np : 1);
    int *core_of = ws.alloc_n<int>(np > 0 ? np : 1);
    unsigned char *enclosed = ws.alloc_n<unsigned char>(ncore > 0 ? ncore : 1);
    float *thr = ws.alloc_n<float>((size_t)(nhalo > 0 ? nhalo : 1) * kMaxShellDivide);
    int *nthr = ws.alloc_n<int>(nhalo > 0 ? nhalo : 1);
    unsigned *kmin = ws.alloc_n<unsigned>(nhalo > 0 ? nhalo : 1);
    unsigned *kmax = ws.alloc_n<unsigned>(nhalo > 0 ? nhalo : 1);
    if (!doff || !dcoff || !dsbase || !dcbase || !dccap || !dnshell || !order || !next || !herr ||
        !pk || !core_of || !enclosed || !thr || !nthr || !kmin || !kmax)
        return hipErrorOutOfMemory;
    PGALF_HIP_TRY(hipMemcpyAsync(doff, off_h, sizeof(long long) * (nhalo + 1),
                                 hipMemcpyHostToDevice, s));
    PGALF_HIP_TRY(hipMemcpyAsync(dcoff, coff_h, sizeof(long long) * (nhalo + 1),
                                 hipMemcpyHostToDevice, s));
    if (nhalo > 0) {
        PGALF_HIP_TRY(hipMemcpyAsync(dsbase, sbase.data(), sizeof(long long) * nhalo,
                                     hipMemcpyHostToDevice, s));
        PGALF_HIP_TRY(hipMemcpyAsync(dcbase, cbase.data(), sizeof(long long) * nhalo,
                                     hipMemcpyHostToDevice, s));
        PGALF_HIP_TRY(hipMemcpyAsync(dccap, ccap.data(), sizeof(long long) * nhalo,
                                     hipMemcpyHostToDevice, s));
        PGALF_HIP_TRY(fill_u32(reinterpret_cast<unsigned *>(herr), nhalo, 0u, s));
        PGALF_HIP_TRY(fill_u64(reinterpret_cast<unsigned long long *>(dnshell), nhalo, 0ull, s));
    }
    if (!small.empty())
        PGALF_HIP_TRY(hipMemcpyAsync(order, small.data(), sizeof(int) * small.size(),
                                     hipMemcpyHostToDevice, s));
    PGALF_HIP_TRY(fill_u32(reinterpret_cast<unsigned *>(next), 2, 0u, s));
    CatArgs a{};
    a.off = doff;
    a.coff = dcoff;
    a.order = order;
    a.ntask = (int)small.size();
    a.next = next;
    a.den = den;
    a.nbr = nbr;
    a.k = k;
    a.type = type;
    a.mass = mass;
    a.peak = peak;
    a.coredensity = o.coredensity;
    a.nummem = o.nummem;
    a.numstar = o.numstar;
    a.starmass = o.starmass;
    a.haloid = o.haloid;
    a.flag = o.flag;
    a.pk = pk;
    a.core_of = core_of;
    a.herr = herr;
    a.thr = thr;
    a.nthr = nthr;
    a.enclosed = enclosed;
    a.part_ptr = so.part_ptr;
    a.core_ptr = so.core_ptr;
    a.sbase = dsbase;
    a.part = so.part;
    a.cores = so.cores;
    a.cbase = dcbase;
    a.ccap = dccap;
    a.nshell = dnshell;
    int slots = 0;
    if (!small.empty()) {
        a.nw = (nmax_small + 31) / 32;
        a.qcap = nmax_small + kBlkT;
        const long long per = 2 * a.nw * 4 + a.qcap * 4;
        long long want = (1LL << 30) / per;
        want = want < 1024 ? want : 1024;
        want = want < (long long)small.size() ? want : (long long)small.size();
        slots = (int)(want > 1 ? want : 1);
        a.bits = ws.alloc_n<unsigned>((size_t)slots * 2 * a.nw);
        a.q = ws.alloc_n<int>((size_t)slots * a.qcap);
        if (!a.bits || !a.q) return hipErrorOutOfMemory;
        PGALF_HIP_TRY(fill_u32(a.bits, (long long)slots * 2 * a.nw, 0u, s));
    }
    std::vector<int> herr_v(nhalo > 0 ? nhalo : 1, 0);
    std::vector<long long> nshell_v(nhalo > 0 ? nhalo : 1, 0);
    // FindCoreDensity, thresholds and shells of halo h by the one-halo tiers
    // on stream hs with scratch from hw (each thread its own)
    auto one_halo = [&](int h, Workspace &hw, hipStream_t hs) -> hipError_t {
        const long long p0 = off_h[h], c0 = coff_h[h], n = off_h[h + 1] - p0;
        const int nc = (int)(coff_h[h + 1] - c0);
        CoreOut oh{o.coredensity + c0, o.nummem + c0, o.numstar + c0, o.starmass + c0,
                   o.haloid + p0, o.flag + p0};
        int fail = 0;
        PGALF_HIP_TRY(find_core_density(den + p0, n, nbr + p0 * k, k, type + p0, mass + p0,
                                        peak + c0, nc, oh, &fail, hw, hs));
        herr_v[h] = fail ? kCatSearch : 0;
        float th[kMaxShellDivide];
        int nt = 0;
        PGALF_HIP_TRY(shell_thresholds(den + p0, n, o.coredensity + c0, nc, th, &nt, hw, hs));
        ShellOut sh{so.part_ptr + sbase[h], so.part + p0, so.core_ptr + sbase[h],
                    so.cores + cbase[h], (long long)nc * nt + 1, ccap[h]};
        long long ns = 0;
        int err = 0;
        PGALF_HIP_TRY(core_shells(den + p0, n, nbr + p0 * k, k, peak + c0, o.coredensity + c0, nc,
                                  th, nt, o.flag + p0, o.haloid + p0, sh, &ns, &err, hw, hs));
        nshell_v[h] = ns;
        if (err && !herr_v[h]) herr_v[h] = 10 + err;
        return hipSuccess;
    };
    // The workgroup halos run on a side stream while this thread and
    // kHelpers helper threads take the large halos (largest first) through
    // the one-halo tiers: the workgroups take one halo each from a queue,
    // and the one-halo tiers' kernels fill the compute units it leaves.  The
    // workgroup thread allocates nothing (its buffers are above); every
    // thread writes only its own halos' outputs and host entries.
    std::vector<int> fcd_err(nhalo > 0 ? nhalo : 1, 0), shell_err(nhalo > 0 ? nhalo : 1, 0);
    std::vector<long long> nshell_small(nhalo > 0 ? nhalo : 1, 0);
    std::stable_sort(big.begin(), big.end(), [&](int a2, int b2) {
        return off_h[a2 + 1] - off_h[a2] > off_h[b2 + 1] - off_h[b2];
    });
    std::atomic<int> next_big(0);
    std::atomic<unsigned> helper_status(0);
    hipError_t side_e = hipSuccess, helper_e[kHelpers];
    for (int i = 0; i < kHelpers; i++) helper_e[i] = hipSuccess;
    std::vector<std::thread> threads;
    struct Joiner {
        std::vector<std::thread> &t;
        ~Joiner() {
            for (auto &x : t)
                if (x.joinable()) x.join();
        }
    } joiner{threads};
    int dev = 0;
    PGALF_HIP_TRY(hipGetDevice(&dev));
    SideSet *ss = nullptr;
    const int nhelp = (int)(big.size() > 1 ? (big.size() - 1 < (size_t)kHelpers ? big.size() - 1
                                                                                 : kHelpers)
                                           : 0);
    if (!small.empty() || nhelp > 0) {
        PGALF_HIP_TRY(side_set(dev, &ss));
        PGALF_HIP_TRY(hipEventRecord(ss->ev, s));
    }
    // The threads copy their work closures, whose references are to objects
    // declared before `joiner` (so they outlive the threads whichever way
    // this function returns); a thread that cannot start leaves its work to
    // this thread.
    bool wg_inline = false;
    const hipStream_t s2 = ss ? ss->wg : nullptr;
    auto wg_work = [&, dev, s2]() -> hipError_t {
                PGALF_HIP_TRY(hipSetDevice(dev));
                cat_fcd_k<<<slots, kBlkT, 0, s2>>>(a);
                PGALF_HIP_TRY(hipGetLastError());
                PGALF_HIP_TRY(to_host(fcd_err.data(), herr, nhalo, s2));
                // thresholds, formed on the host as the reference forms them
                cat_minmax_k<<<(unsigned)small.size(), 256, 0, s2>>>(order, doff, dcoff, den,
                                                                     o.coredensity, kmin, kmax);
                std::vector<unsigned> lo(nhalo), hi(nhalo);
                PGALF_HIP_TRY(to_host(lo.data(), kmin, nhalo, s2));
                PGALF_HIP_TRY(to_host(hi.data(), kmax, nhalo, s2));
                std::vector<float> th((size_t)nhalo * kMaxShellDivide, 0.0f);
                std::vector<int> nt(nhalo, 0), mark(nhalo, 1);
                for (int h : small) {
                    const long long n = off_h[h + 1] - off_h[h], nc = coff_h[h + 1] - coff_h[h];
                    const float pmin = host_val(lo[h]);
                    const float cmax = nc > 0 ? host_val(hi[h]) : -1.e27f;
                    nt[h] = thresholds_from(pmin, cmax, n, th.data() + (size_t)h * kMaxShellDivide);
                    mark[h] = fcd_err[h] ? 1 : 0;  // halos marked 1 are skipped
                }
                PGALF_HIP_TRY(hipMemcpyAsync(thr, th.data(), sizeof(float) * th.size(),
                                             hipMemcpyHostToDevice, s2));
                PGALF_HIP_TRY(hipMemcpyAsync(nthr, nt.data(), sizeof(int) * nhalo,
                                             hipMemcpyHostToDevice, s2));
                PGALF_HIP_TRY(hipMemcpyAsync(herr, mark.data(), sizeof(int) * nhalo,
                                             hipMemcpyHostToDevice, s2));
                PGALF_HIP_TRY(fill_u32(reinterpret_cast<unsigned *>(next), 1, 0u, s2));
                cat_shells_k<<<slots, kBlkT, 0, s2>>>(a);
                PGALF_HIP_TRY(hipGetLastError());
                PGALF_HIP_TRY(to_host(shell_err.data(), herr, nhalo, s2));
                return to_host(nshell_small.data(), dnshell, nhalo, s2);
    };
    if (!small.empty()) {
        PGALF_HIP_TRY(hipStreamWaitEvent(s2, ss->ev, 0));
        try {
            threads.emplace_back([&side_e, wg_work]() { side_e = wg_work(); });
        } catch (const std::system_error &) {
            wg_inline = true;
        }
    }
    for (int i = 0; i < nhelp; i++) {
        const hipStream_t hs = ss->s[i];
        Workspace *hw = ss->ws[i];
        PGALF_HIP_TRY(hipStreamWaitEvent(hs, ss->ev, 0));
        auto helper_work = [&, dev, hs, hw]() -> hipError_t {
                PGALF_HIP_TRY(hipSetDevice(dev));
                WsCall call(*hw, hs);
                PGALF_HIP_TRY(call.e);
                for (int j; (j = next_big.fetch_add(1)) < (int)big.size();)
                    PGALF_HIP_TRY(one_halo(big[j], *hw, hs));
                // the helper's device error bits go to the caller's status
                unsigned st = 0;
                PGALF_HIP_TRY(to_host(&st, hw->status(), 1, hs));
                if (st) {
                    helper_status.fetch_or(st);
                    PGALF_HIP_TRY(fill_u32(hw->status(), 1, 0u, hs));
                }
                return hipStreamSynchronize(hs);
        };
        try {
            threads.emplace_back([&helper_e, i, helper_work]() { helper_e[i] = helper_work(); });
        } catch (const std::system_error &) {
            break;  // the large halos left go to this thread
        }
    }
    for (int j; (j = next_big.fetch_add(1)) < (int)big.size();)
        PGALF_HIP_TRY(one_halo(big[j], ws, s));
    if (wg_inline) side_e = wg_work();
    for (auto &x : threads) x.join();
    PGALF_HIP_TRY(side_e);
    for (int i = 0; i < kHelpers; i++) PGALF_HIP_TRY(helper_e[i]);
    if (helper_status.load()) {
        or_status_k<<<1, 1, 0, s>>>(ws.status(), helper_status.load());
        PGALF_HIP_TRY(hipGetLastError());
    }
    for (int h : small) {
        if (fcd_err[h] || shell_err[h]) {
            // a sweep or the core entries outgrew the workgroup path: the
            // halo again through the one-halo tiers
            PGALF_HIP_TRY(one_halo(h, ws, s));
        } else {
            nshell_v[h] = nshell_small[h];
        }
    }
    for (int h = 0; h < nhalo; h++) {
        herr_h[h] = herr_v[h];
        nshell_h[h] = nshell_v[h];
    }
    return hipGetLastError();
}

}  // namespace pgalf
