// Device-wide exclusive scan and stable LSD radix sort for gfx950 (wave64).
//
// Scan: reduce-then-scan over tiles of 2048 uint32 (256 threads x 8).
// Sort: 8-bit digits; per pass a tile histogram, a scan of the digit-major
// count table, and a stable scatter that ranks equal digits in element order
// with eight 64-lane ballots per round (no atomics in the ranking, so the
// output order is fully determined by the input order).
#include "pgalf_internal.h"

namespace pgalf {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;  // 2048
constexpr int kWaves = kThreads / 64;
constexpr long long kSmallScan = 8LL * kTile;  // scanned by one block in one launch

__device__ inline unsigned wave_incl_scan(unsigned v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        unsigned t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
__device__ inline unsigned block_excl_scan(unsigned v, unsigned *total) {
    __shared__ unsigned wsum[kWaves];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned inc = wave_incl_scan(v, lane);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    unsigned off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; i++) {
        unsigned s = wsum[i];
        off += (i < w) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ void __launch_bounds__(kThreads) scan_reduce_k(const unsigned *in, long long n,
                                                          unsigned *bsum) {
    long long base = (long long)blockIdx.x * kTile;
    unsigned s = 0;
#pragma unroll
    for (int r = 0; r < kItems; r++) {
        long long i = base + r * kThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    unsigned tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Single block: exclusive scan of nb values, chunk by chunk (in may equal
// out: each thread reads its values before writing them).  Scans the block
// sums of a multi-block scan, and short arrays whole.
__global__ void __launch_bounds__(kThreads) scan_top_k(const unsigned *in, unsigned *out,
                                                       long long nb, unsigned *total) {
    unsigned carry = 0;
    for (long long c0 = 0; c0 < nb; c0 += kTile) {
        unsigned v[kItems];
        unsigned s = 0;
#pragma unroll
        for (int r = 0; r < kItems; r++) {
            long long i = c0 + (long long)threadIdx.x * kItems + r;
            v[r] = i < nb ? in[i] : 0u;
            s += v[r];
        }
        unsigned tot;
        unsigned off = block_excl_scan(s, &tot) + carry;
#pragma unroll
        for (int r = 0; r < kItems; r++) {
            long long i = c0 + (long long)threadIdx.x * kItems + r;
            if (i < nb) out[i] = off;
            off += v[r];
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(kThreads) scan_apply_k(const unsigned *in, unsigned *out,
                                                         long long n, const unsigned *bsum) {
    __shared__ unsigned tile[kTile];
    long long base = (long long)blockIdx.x * kTile;
#pragma unroll
    for (int r = 0; r < kItems; r++) {
        long long i = base + r * kThreads + threadIdx.x;
        tile[r * kThreads + threadIdx.x] = i < n ? in[i] : 0u;
    }
    __syncthreads();
    unsigned v[kItems];
    unsigned s = 0;
#pragma unroll
    for (int r = 0; r < kItems; r++) {
        v[r] = tile[threadIdx.x * kItems + r];
        s += v[r];
    }
    unsigned tot;
    unsigned off = block_excl_scan(s, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int r = 0; r < kItems; r++) {
        tile[threadIdx.x * kItems + r] = off;
        off += v[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kItems; r++) {
        long long i = base + r * kThreads + threadIdx.x;
        if (i < n) out[i] = tile[r * kThreads + threadIdx.x];
    }
}

// ---- radix sort ------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) rs_hist_k(const unsigned *keys, long long n,
                                                      int shift, long long nb,
                                                      unsigned *counts) {
    __shared__ unsigned hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    long long base = (long long)blockIdx.x * kTile;
#pragma unroll
    for (int r = 0; r < kItems; r++) {
        long long i = base + r * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(long long)threadIdx.x * nb + blockIdx.x] = hist[threadIdx.x];
}

// One LSD pass over the block's tile: every item is ranked within its digit
// (stable: by round, wave and lane), placed in LDS in digit order, and the
// tile is then written out digit run by digit run, so consecutive threads
// store consecutive positions of one run.
__global__ void __launch_bounds__(kThreads) rs_scatter_k(const unsigned *keys,
                                                         const unsigned *vals, long long n,
                                                         int shift, long long nb,
                                                         const unsigned *offsets,
                                                         unsigned *okeys, unsigned *ovals) {
    __shared__ unsigned s_key[kTile], s_val[kTile];
    __shared__ unsigned gbase[256];
    __shared__ unsigned lstart[256];
    __shared__ unsigned run[256];
    __shared__ unsigned wcnt[kWaves][256];
    __shared__ unsigned wsum[kWaves];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long long base = (long long)blockIdx.x * kTile;
    const int tcount = (int)(n - base < kTile ? n - base : kTile);
    // this block's count of digit tid from the scanned (digit-major) counts
    const long long ci = (long long)tid * nb + blockIdx.x;
    const unsigned gb = offsets[ci];
    const unsigned cnt = (ci + 1 < 256 * nb ? offsets[ci + 1] : (unsigned)n) - gb;
    gbase[tid] = gb;
    // exclusive scan of the block's digit counts: local start of each run
    unsigned incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    run[tid] = 0;
#pragma unroll
    for (int i = 0; i < kWaves; i++) wcnt[i][tid] = 0;
    __syncthreads();
    unsigned wbase = 0;
    for (int i = 0; i < w; i++) wbase += wsum[i];
    lstart[tid] = wbase + incl - cnt;
    __syncthreads();
    const unsigned long long lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int r = 0; r < kItems; r++) {
        const int li = r * kThreads + tid;
        const bool valid = li < tcount;
        const unsigned key = valid ? keys[base + li] : 0u;
        const unsigned val = valid ? vals[base + li] : 0u;
        const unsigned d = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int bt = 0; bt < 8; bt++) {
            const unsigned long long bb = __ballot((d >> bt) & 1u);
            peers &= ((d >> bt) & 1u) ? bb : ~bb;
        }
        const unsigned rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned off = lstart[d] + run[d] + rank;
            for (int ww = 0; ww < w; ww++) off += wcnt[ww][d];
            s_key[off] = key;
            s_val[off] = val;
        }
        __syncthreads();
        unsigned add = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ww++) {
            add += wcnt[ww][tid];
            wcnt[ww][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
    for (int i = tid; i < tcount; i += kThreads) {
        const unsigned key = s_key[i];
        const unsigned d = (key >> shift) & 255u;
        const unsigned dst = gbase[d] + (unsigned)i - lstart[d];
        okeys[dst] = key;
        ovals[dst] = s_val[i];
    }
}

template <class T>
__global__ void fill_k(T *p, long long n, T v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

template <class T>
hipError_t fill_any(T *p, long long n, T v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    long long nb = (n + kThreads - 1) / kThreads;
    if (nb > 4096) nb = 4096;
    fill_k<T><<<dim3((unsigned)nb), dim3(kThreads), 0, s>>>(p, n, v);
    return hipGetLastError();
}

}  // namespace

hipError_t fill_u32(unsigned *p, long long n, unsigned v, hipStream_t s) {
    return fill_any<unsigned>(p, n, v, s);
}
hipError_t fill_u64(unsigned long long *p, long long n, unsigned long long v, hipStream_t s) {
    return fill_any<unsigned long long>(p, n, v, s);
}

hipError_t exclusive_scan_u32(const unsigned *in, unsigned *out, long long n,
                              unsigned *total, Workspace &ws, hipStream_t s) {
    if (n <= 0) return total ? fill_u32(total, 1, 0u, s) : hipSuccess;
    if (n <= kSmallScan) {  // one launch
        scan_top_k<<<1, kThreads, 0, s>>>(in, out, n, total);
        return hipGetLastError();
    }
    long long nb = (n + kTile - 1) / kTile;
    unsigned *bsum = ws.alloc_n<unsigned>((size_t)nb);
    if (!bsum) return hipErrorOutOfMemory;
    scan_reduce_k<<<dim3((unsigned)nb), dim3(kThreads), 0, s>>>(in, n, bsum);
    scan_top_k<<<1, kThreads, 0, s>>>(bsum, bsum, nb, total);
    scan_apply_k<<<dim3((unsigned)nb), dim3(kThreads), 0, s>>>(in, out, n, bsum);
    return hipGetLastError();
}

hipError_t radix_sort_pairs(unsigned **keys, unsigned **vals, unsigned *keys_tmp,
                            unsigned *vals_tmp, long long n, int bits, Workspace &ws,
                            hipStream_t s) {
    if (n <= 1) return hipSuccess;
    long long nb = (n + kTile - 1) / kTile;
    unsigned *counts = ws.alloc_n<unsigned>((size_t)(nb * 256));
    if (!counts) return hipErrorOutOfMemory;
    unsigned *ka = *keys, *va = *vals, *kb = keys_tmp, *vb = vals_tmp;
    for (int shift = 0; shift < bits; shift += 8) {
        rs_hist_k<<<dim3((unsigned)nb), dim3(kThreads), 0, s>>>(ka, n, shift, nb, counts);
        PGALF_HIP_TRY(exclusive_scan_u32(counts, counts, nb * 256, nullptr, ws, s));
        rs_scatter_k<<<dim3((unsigned)nb), dim3(kThreads), 0, s>>>(ka, va, n, shift, nb,
                                                                    counts, kb, vb);
        unsigned *t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    *keys = ka;
    *vals = va;
    return hipGetLastError();
}

}  // namespace pgalf
