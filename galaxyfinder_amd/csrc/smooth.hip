// FFT Gaussian smoothing for gfx950 (gaussian_Smoothing,
// NewGalFinder/gsmooth.c:14-115).
//
// The reference transforms the padded grid with a 3-D r2c FFT, multiplies
// bin (i,j,k) by fx(i)*fy(j)*fz(k), fd(k) = exp(-(pi RG_c)^2 (k'/n_d)^2)
// with k' the signed frequency, and transforms back with a 1/(nx ny nz)
// scale.  Because the multiplier is a product of real, even, per-axis
// factors, the operator is exactly three 1-D circular convolutions, one per
// axis.  Each axis pass reads the grid once and writes it once, two real
// lines packed as one complex line, with one of three kernels:
//   * lines of 64, 128, 256: a four-step FFT with each thread's DFT-8/16 in
//     FP64 registers and two LDS transposes per line (smooth_fft2_k);
//   * short lines (up to the direct threshold, 56 by default): direct
//     circular convolution with the periodic kernel (smooth_direct_k,
//     smooth_direct_x_k for rows; small cubes in one LDS-resident kernel);
//   * other lengths: a workgroup loads a batch of lines into LDS, runs a
//     Stockham radix-4/2 FFT, multiplies by the real spectral factor H, runs
//     the inverse FFT and stores the lines back (smooth_axis_k).
// FFT lengths:
//   n_d a power of two:  L = n_d and H(k) = fd(k) / n_d.
//   any other n_d:       L = pow2 >= 2 n_d - 1, the line is zero-padded and
//                        H = DFT_L(h) / L with h the periodic kernel g_d laid
//                        out circularly over [-(n_d-1), n_d-1]; the first n_d
//                        outputs of the length-L linear convolution are the
//                        length-n_d circular one.
// The grid stays float32 in HBM; each line is transformed in FP64 (FP64
// vector FMA runs at the FP32 rate on this part and the passes are
// memory-bound), so the only FP32 roundings are the three stores: every
// cell keeps ~1 ulp relative accuracy even far below the line's peak, which
// an FP32 FFT (whose error scales with the line norm) cannot give.
// H and the twiddles are formed on the host in FP64.
// Pad columns [nx, mx) are zeroed by the x pass (the reference leaves FFTW
// scratch there).
#include <cmath>
#include <cstring>
#include <array>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "pgalf_internal.h"

namespace pgalf {
namespace {

constexpr int kBlock = 256;
// Complex FP64 elements of line data per block (P * L <= LC): it sizes the
// per-thread butterfly registers of a pass.  LC = 2048 keeps the kernel at
// ~60 VGPRs; the longest lines (L = 4096) use the LC = 4096 instance.
constexpr int kLdsComplex = 2048;
constexpr int kLdsComplexMax = 4096;
constexpr int kMaxL = 4096;

struct AxisPass {
    int n;              // line length (grid points)
    int L;              // FFT length (power of two)
    long long sa;       // stride along the line
    int nb;             // number of lines along the fast batch dimension
    long long sb;       // stride of that dimension
    long long sc;       // stride of the slow batch dimension (one per blockIdx.y)
    int P;              // complex lines per block (2P real lines)
    int contiguous;     // 1: lines are contiguous (x pass)
    int zero_pads;      // x pass: zero columns [n, padw) of each row
    int padw;           // row width for the pad zeroing (mx)
    int LS;             // line stride in LDS (complex elements)
    int tables;         // 1: twiddles and spectral factor staged in LDS
    int pairs;          // 1: adjacent lines may be read as one float2 (8-byte aligned)
    long long gcells;   // floats of one grid (register FFT: byte offsets fit 31 bits)
};

// One grid's lines in a pass that covers grids of several shapes (a batch's
// axis pass over every grid with the same line length): its offset, the
// stride along the line, the fast batch dimension (nb lines, stride sb) and
// the slow one (nc, stride sc).  Blocks beyond a grid's own lines exit.
struct GridLines {
    long long goff;
    long long sa, sb, sc;
    int nb, nc;
};

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// One Stockham pass over P lines of length L in LDS (in place: all inputs of
// this thread are read before the barrier, outputs written after it).
template <int R, int LC>
__device__ inline void stockham_pass(double2 *buf, int L, int LS, int P, int Ns,
                                     const double2 *tw, bool inverse) {
    // at most LC/R butterflies per pass -> kMaxBf per thread; the
    // per-thread loop is fully unrolled so v[][] stays in registers.
    constexpr int kMaxBf = LC / R / kBlock;
    const int nbf_line = L / R;
    const int nbf = P * nbf_line;
    const int tstep = L / (Ns * R);  // twiddle index step
    double2 v[kMaxBf][R];
    int dst[kMaxBf];
#pragma unroll
    for (int c = 0; c < kMaxBf; c++) {
        int b = threadIdx.x + c * kBlock;
        if (b >= nbf) continue;
        int p = b / nbf_line, j = b - p * nbf_line;
        double2 *line = buf + p * LS;
        int k = j % Ns;
#pragma unroll
        for (int r = 0; r < R; r++) {
            double2 a = line[j + r * nbf_line];
            if (r) {
                double2 w = tw[(r * k * tstep) & (L - 1)];
                if (inverse) w.y = -w.y;
                a = cmul(a, w);
            }
            v[c][r] = a;
        }
        if (R == 2) {
            double2 a0 = v[c][0], a1 = v[c][1];
            v[c][0] = make_double2(a0.x + a1.x, a0.y + a1.y);
            v[c][1] = make_double2(a0.x - a1.x, a0.y - a1.y);
        } else {
            double2 a0 = v[c][0], a1 = v[c][1], a2 = v[c][2], a3 = v[c][3];
            double2 s02 = make_double2(a0.x + a2.x, a0.y + a2.y);
            double2 d02 = make_double2(a0.x - a2.x, a0.y - a2.y);
            double2 s13 = make_double2(a1.x + a3.x, a1.y + a3.y);
            double2 d13 = make_double2(a1.x - a3.x, a1.y - a3.y);
            // forward twiddle W4 = -i: (-i)(x + iy) = y - ix; inverse +i
            double2 md13 = inverse ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);
            v[c][0] = make_double2(s02.x + s13.x, s02.y + s13.y);
            v[c][1] = make_double2(d02.x + md13.x, d02.y + md13.y);
            v[c][2] = make_double2(s02.x - s13.x, s02.y - s13.y);
            v[c][3] = make_double2(d02.x - md13.x, d02.y - md13.y);
        }
        dst[c] = p * LS + (j / Ns) * Ns * R + k;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kMaxBf; c++) {
        int b = threadIdx.x + c * kBlock;
        if (b >= nbf) continue;
#pragma unroll
        for (int r = 0; r < R; r++) buf[dst[c] + r * Ns] = v[c][r];
    }
    __syncthreads();
}

template <int LC>
__device__ inline void fft_lines(double2 *buf, int L, int LS, int P, const double2 *tw,
                                 bool inverse) {
    int Ns = 1;
    while (Ns * 4 <= L) {
        stockham_pass<4, LC>(buf, L, LS, P, Ns, tw, inverse);
        Ns *= 4;
    }
    if (Ns < L) stockham_pass<2, LC>(buf, L, LS, P, Ns, tw, inverse);
}

// LDS per block: line data (P lines of LS complex), plus the twiddles and
// the spectral factor when L <= kLdsTables.  LS = L + 1: one complex element
// of padding per line, so the strided (y/z) passes, whose consecutive
// threads write consecutive lines, hit different banks.
constexpr int kLdsBytes = 65536;
constexpr int kLdsTables = 1024;
constexpr int kLoadBatch = 16;  // global loads in flight per thread in the load phase

// blockIdx.x: batch of 2P lines; blockIdx.y: slow batch index; blockIdx.z:
// grid of a batch of same-shape grids at offsets goffs (nullptr: one grid).
template <int LC>
__global__ void __launch_bounds__(kBlock)
smooth_axis_k(float *g, AxisPass ap, const double2 *tw_g, const double *H_g,
              const long long *goffs, const GridLines *gl) {
    extern __shared__ double2 buf[];
    const int L = ap.L, P = ap.P, n = ap.n, LS = ap.LS;
    const long long ib0 = (long long)blockIdx.x * 2 * P;
    long long nbl = ap.nb, sbl = ap.sb, sal = ap.sa;
    float *base;
    if (gl) {
        const GridLines d = gl[blockIdx.z];
        if ((int)blockIdx.y >= d.nc || ib0 >= d.nb) return;  // the whole block
        nbl = d.nb;
        sbl = d.sb;
        sal = d.sa;
        base = g + d.goff + (long long)blockIdx.y * d.sc;
    } else {
        base = g + (goffs ? goffs[blockIdx.z] : 0LL) + (long long)blockIdx.y * ap.sc;
    }
    const bool tables = ap.tables != 0;
    double2 *s_tw = buf + P * LS;
    double *s_H = reinterpret_cast<double *>(s_tw + L);
    if (tables) {
        for (int i = threadIdx.x; i < L; i += kBlock) {
            s_tw[i] = tw_g[i];
            s_H[i] = H_g[i];
        }
    }
    const double2 *tw = tables ? s_tw : tw_g;
    const double *H = tables ? s_H : H_g;
    const int lines = 2 * P;
    // load: line lb (0..2P-1) element j -> buf[(lb/2)*LS + j].(x|y); kLoadBatch
    // independent loads per thread are issued before their LDS stores so
    // their latencies overlap
    const int tot = lines * n;
    for (int i0 = threadIdx.x; i0 < tot; i0 += kBlock * kLoadBatch) {
        float v[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int idx = i0 + u * kBlock;
            int lb = 0, j = 0;
            if (ap.contiguous) { lb = idx / n; j = idx - lb * n; }
            else { j = idx / lines; lb = idx - j * lines; }
            const long long ib = ib0 + lb;
            v[u] = (idx < tot && ib < nbl) ? base[ib * sbl + (long long)j * sal] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int idx = i0 + u * kBlock;
            if (idx < tot) {
                int lb, j;
                if (ap.contiguous) { lb = idx / n; j = idx - lb * n; }
                else { j = idx / lines; lb = idx - j * lines; }
                double *dst = reinterpret_cast<double *>(&buf[(lb >> 1) * LS + j]);
                dst[lb & 1] = (double)v[u];
            }
        }
    }
    for (int idx = threadIdx.x; idx < P * (L - n); idx += kBlock) {
        int p = idx / (L - n), j = n + idx % (L - n);
        buf[p * LS + j] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lines<LC>(buf, L, LS, P, tw, false);
    for (int idx = threadIdx.x; idx < P * L; idx += kBlock) {
        const int p = idx / L, j = idx - p * L;
        double h = H[j];
        double2 a = buf[p * LS + j];
        buf[p * LS + j] = make_double2(a.x * h, a.y * h);
    }
    __syncthreads();
    fft_lines<LC>(buf, L, LS, P, tw, true);
    for (int idx = threadIdx.x; idx < tot; idx += kBlock) {
        int lb, j;
        if (ap.contiguous) { lb = idx / n; j = idx - lb * n; }
        else { j = idx / lines; lb = idx - j * lines; }
        long long ib = ib0 + lb;
        if (ib < nbl) {
            const double *src = reinterpret_cast<const double *>(&buf[(lb >> 1) * LS + j]);
            base[ib * sbl + (long long)j * sal] = (float)src[lb & 1];
        }
    }
    if (ap.zero_pads) {
        const int pw = ap.padw - n;
        for (int idx = threadIdx.x; idx < lines * pw; idx += kBlock) {
            int lb = idx / pw, j = n + idx % pw;
            long long ib = ib0 + lb;
            if (ib < nbl) base[ib * sbl + j] = 0.0f;
        }
    }
}

// ---- register-resident four-step FFT for lines of 64, 128 and 256 -------------
// L = RA * RB (64 = 8 x 8, 128 = 8 x 16, 256 = 16 x 16).  With j = j1 + RB j2
// and k = k2 + RA k1,
//   X[k2 + RA k1] = sum_j1 W_RB^{j1 k1} W_L^{j1 k2} sum_j2 W_RA^{j2 k2} x[j1 + RB j2].
// Step 1: thread j1 loads x[j1 + RB j2] (j2 < RA) of its two real lines
// straight from the grid, runs a DFT_RA in registers and applies W_L^{j1 k2};
// an LDS transpose gives thread k2 the row Y[k2][j1] (j1 < RB).  Step 2:
// DFT_RB -> X[k2 + RA k1], times the real factor H, inverse DFT_RB and
// W_L^{-j1 k2}; the inverse transform keeps the same element order, so
// nothing moves between the two.  A second transpose returns column j1 to
// thread j1, whose inverse DFT_RA yields x'[j1 + RB j2] -- the elements it
// loaded -- and it stores them straight back.  Two LDS round trips per line
// in all, every DFT in FP64 registers, one FP32 rounding per output.

// cos and sin of 2 pi m / 16 (m after unrolling is a constant)
__device__ constexpr double kCos1 = 0x1.d906bcf328d46p-1;  // cos(pi/8)
__device__ constexpr double kCos2 = 0x1.6a09e667f3bcdp-1;  // cos(pi/4)
__device__ constexpr double kCos3 = 0x1.87de2a6aea963p-2;  // cos(3pi/8)
__device__ constexpr double cos16(int m) {
    return m == 0 ? 1.0 : m == 1 ? kCos1 : m == 2 ? kCos2 : m == 3 ? kCos3 : m == 4 ? 0.0
         : m == 5 ? -kCos3 : m == 6 ? -kCos2 : -kCos1;  // m < 8
}
__device__ constexpr double sin16(int m) {
    return m == 0 ? 0.0 : m == 1 ? kCos3 : m == 2 ? kCos2 : m == 3 ? kCos1 : m == 4 ? 1.0
         : m == 5 ? kCos1 : m == 6 ? kCos2 : kCos3;  // m < 8
}
// b * exp(-+2 pi i m / 16): forward takes the minus sign, INV the plus
// FMA contraction is allowed in this FFT (it follows no reference rounding;
// the parity bar is against the FP64 smoothing).
__device__ inline double2 cmul_c(double2 a, double2 b) {
#pragma clang fp contract(fast)
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <bool INV>
__device__ inline double2 tw16(double2 b, int m) {
#pragma clang fp contract(fast)
    if (m == 0) return b;
    if (m == 4) return INV ? make_double2(-b.y, b.x) : make_double2(b.y, -b.x);
    const double c = cos16(m), s = INV ? sin16(m) : -sin16(m);
    return make_double2(b.x * c - b.y * s, b.x * s + b.y * c);
}
template <int R>
__device__ constexpr int bitrev(int i) {
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}
// radix-2 decimation-in-time stages (SPAN = 1, 2, ..., R/2)
template <int R, bool INV, int SPAN>
__device__ inline void dit_stages(double2 (&v)[R]) {
    if constexpr (SPAN < R) {
#pragma unroll
        for (int s0 = 0; s0 < R; s0 += 2 * SPAN) {
#pragma unroll
            for (int k = 0; k < SPAN; k++) {
                const double2 a = v[s0 + k];
                const double2 b = tw16<INV>(v[s0 + k + SPAN], k * (16 / (2 * SPAN)));
                v[s0 + k] = make_double2(a.x + b.x, a.y + b.y);
                v[s0 + k + SPAN] = make_double2(a.x - b.x, a.y - b.y);
            }
        }
        dit_stages<R, INV, 2 * SPAN>(v);
    }
}
// In-register DFT of size R <= 16, natural order in and out (unscaled).
template <int R, bool INV>
__device__ inline void dft_reg(double2 (&v)[R]) {
    static_assert(R >= 2 && R <= 16 && (R & (R - 1)) == 0, "radix");
#pragma unroll
    for (int i = 0; i < R; i++) {
        const int j = bitrev<R>(i);
        if (j > i) {
            const double2 t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
    dit_stages<R, INV, 1>(v);
}
// Twiddle bases of one thread: p[b] = W_L^{b u} (b = 1..3) and
// q[a] = W_L^{4 a u} (a = 1..RQ) from the table tw[m] = exp(-2 pi i m / L).
template <int RQ>
struct TwBase {
    double2 p[3], q[RQ];
};
template <int RQ>
__device__ inline TwBase<RQ> twiddle_base(const double2 *tw, int u, int L) {
    TwBase<RQ> t;
#pragma unroll
    for (int b = 1; b < 4; b++) t.p[b - 1] = tw[(b * u) & (L - 1)];
#pragma unroll
    for (int a = 1; a <= RQ; a++) t.q[a - 1] = tw[(4 * a * u) & (L - 1)];
    return t;
}
// v[k] *= W_L^{+-u k} (k < R): the powers 1..3 and the multiples of 4 are
// table entries, the rest one product of two.
template <int R, bool INV, int RQ>
__device__ inline void twiddle_row(double2 (&v)[R], const TwBase<RQ> &t) {
    static_assert(R % 4 == 0 && R / 4 - 1 <= RQ, "whole quads");
#pragma unroll
    for (int k = 1; k < R; k++) {
        const int a = k / 4, b = k % 4;
        double2 w = a == 0 ? t.p[b - 1] : b == 0 ? t.q[a - 1] : cmul_c(t.q[a - 1], t.p[b - 1]);
        if (INV) w.y = -w.y;
        v[k] = cmul_c(v[k], w);
    }
}

// Line groups: one group is NL complex lines (2 NL real lines, consecutive
// in the batch dimension) at one slow index of one grid; RB threads per
// line.  A block walks the groups with a grid stride and loads the next
// group's elements into registers before transforming the current one, so
// the memory latency overlaps the arithmetic.  CONTIG (rows): lanes run along
// the line, so loads are coalesced within a row; otherwise lanes run across
// lines (consecutive x) and a line pair is one float2.  LDS rows are padded
// (RB + 1 per k2 row, odd line stride) so both transposes spread over the
// banks.
struct Fft2Groups {
    int nbx;           // groups along the batch dimension
    int nc;            // slow batch indices
    unsigned ngroups;  // nbx * nc * grids (< 2^31)
};

// Grid memory is read and written with buffer instructions: a wave-uniform
// descriptor per group (its slice of the grid), one 32-bit byte offset per
// lane and a uniform offset per element, so no per-element 64-bit address
// is held in registers.
typedef unsigned int u32x2 __attribute__((__vector_size__(8)));
constexpr int kRsrcWord3 = 0x00020000;  // raw dword access (gfx9 family)

template <int RA, int RB, int NL, bool CONTIG>
__global__ void __launch_bounds__(NL * RB) __attribute__((amdgpu_waves_per_eu(2)))
smooth_fft2_k(float *g, AxisPass ap, const double2 *tw, const double *H,
              const long long *goffs, Fft2Groups gr) {
    static_assert(RB >= RA, "step-1 threads cover every line element");
    constexpr int L = RA * RB;
    constexpr int NT = NL * RB;
    constexpr int ROW = RB + 1;
    constexpr int LSTR = (RA * ROW) | 1;
    constexpr int RQ = RB / 4 - 1;
    __shared__ double2 buf[NL * LSTR];
    __shared__ double s_H[L];
    const int tid = threadIdx.x;
    const int line = CONTIG ? tid / RB : tid % NL;
    const int u = CONTIG ? tid % RB : tid / NL;
    for (int i = tid; i < L; i += NT) s_H[i] = H[i];
    const TwBase<RQ> twb = twiddle_base<RQ>(tw, u, L);
    double2 *row = buf + line * LSTR;
    // byte strides: along the line (per element j2 of this thread) and
    // between the two lines of a pair
    const int step = RB * (int)ap.sa * 4;
    const int bstep = (int)ap.sb * 4;

    // group gi: its slice of the grid (wave-uniform) and this lane's lines
    struct Grp {
        float *base;  // the group's slice
        int bytes;    // bytes from base to the end of the grid
        int off;      // byte offset of element u of line a
        bool va, vb;  // lines a, b present
    };
    auto locate = [&](unsigned gi) {
        Grp q;
        const unsigned rest = gi / (unsigned)gr.nbx;
        const unsigned bx = gi - rest * (unsigned)gr.nbx;
        const unsigned c = rest % (unsigned)gr.nc, z = rest / (unsigned)gr.nc;
        const long long c0 = (long long)c * ap.sc;
        q.base = g + (goffs ? goffs[z] : 0LL) + c0;
        q.bytes = (int)((ap.gcells - c0) * 4);
        const int la = (int)bx * 2 * NL + 2 * line;
        q.va = la < ap.nb;
        q.vb = la + 1 < ap.nb;
        q.off = (int)(((long long)la * ap.sb + (long long)u * ap.sa) * 4);
        return q;
    };
    auto rsrc = [](const Grp &q) {
        return __builtin_amdgcn_make_buffer_rsrc(q.base, (short)0, q.bytes, kRsrcWord3);
    };
    auto load = [&](const Grp &q, float2 (&x)[RA]) {
        const auto r = rsrc(q);
        const bool pair = !CONTIG && ap.pairs && q.vb;
#pragma unroll
        for (int j2 = 0; j2 < RA; j2++) {
            float a = 0.0f, b = 0.0f;
            if (pair) {
                const u32x2 t = __builtin_amdgcn_raw_buffer_load_b64(r, q.off, j2 * step, 0);
                a = __uint_as_float(t[0]);
                b = __uint_as_float(t[1]);
            } else {
                if (q.va) a = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, q.off, j2 * step, 0));
                if (q.vb) b = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, q.off, j2 * step + bstep, 0));
            }
            x[j2] = make_float2(a, b);
        }
    };

    unsigned gi = blockIdx.x;
    Grp cg{};
    float2 cur[RA];
    if (gi < gr.ngroups) {
        cg = locate(gi);
        load(cg, cur);
    }
    __syncthreads();  // s_H
    for (; gi < gr.ngroups; gi += gridDim.x) {
        // next group's elements in flight during this group's transform
        const unsigned gn = gi + gridDim.x;
        Grp ng{};
        float2 nxt[RA];
        if (gn < gr.ngroups) {
            ng = locate(gn);
            load(ng, nxt);
        }
        // step 1: DFT_RA over j2 and W_L^{j1 k2}, j1 = u
        double2 v[RA];
#pragma unroll
        for (int j2 = 0; j2 < RA; j2++) v[j2] = make_double2((double)cur[j2].x, (double)cur[j2].y);
        dft_reg<RA, false>(v);
        twiddle_row<RA, false>(v, twb);
        __syncthreads();  // the previous group's last reads of buf are done
#pragma unroll
        for (int k2 = 0; k2 < RA; k2++) row[k2 * ROW + u] = v[k2];
        __syncthreads();
        // step 2 (k2 = u): each thread reads and rewrites only its own row
        if (u < RA) {
            double2 w[RB];
#pragma unroll
            for (int j1 = 0; j1 < RB; j1++) w[j1] = row[u * ROW + j1];
            dft_reg<RB, false>(w);
#pragma unroll
            for (int k1 = 0; k1 < RB; k1++) {
                const double h = s_H[u + RA * k1];
                w[k1] = make_double2(w[k1].x * h, w[k1].y * h);
            }
            dft_reg<RB, true>(w);
            twiddle_row<RB, true>(w, twb);
#pragma unroll
            for (int j1 = 0; j1 < RB; j1++) row[u * ROW + j1] = w[j1];
        }
        __syncthreads();
        // step 3: inverse DFT_RA over k2 gives x'[u + RB j2]
#pragma unroll
        for (int k2 = 0; k2 < RA; k2++) v[k2] = row[k2 * ROW + u];
        dft_reg<RA, true>(v);
        const bool pair = !CONTIG && ap.pairs && cg.vb;
        const auto cr = rsrc(cg);
#pragma unroll
        for (int j2 = 0; j2 < RA; j2++) {
            const unsigned a = __float_as_uint((float)v[j2].x);
            const unsigned b = __float_as_uint((float)v[j2].y);
            if (pair) {
                u32x2 t;
                t[0] = a;
                t[1] = b;
                __builtin_amdgcn_raw_buffer_store_b64(t, cr, cg.off, j2 * step, 0);
            } else {
                if (cg.va) __builtin_amdgcn_raw_buffer_store_b32(a, cr, cg.off, j2 * step, 0);
                if (cg.vb) __builtin_amdgcn_raw_buffer_store_b32(b, cr, cg.off, j2 * step + bstep, 0);
            }
        }
        if (ap.zero_pads && u == 0) {
            // rows: element j of line a at off + 4 j (u = 0)
            for (int j = ap.n; j < ap.padw; j++) {
                if (cg.va) __builtin_amdgcn_raw_buffer_store_b32(0u, cr, cg.off, 4 * j, 0);
                if (cg.vb) __builtin_amdgcn_raw_buffer_store_b32(0u, cr, cg.off, 4 * j + bstep, 0);
            }
        }
        cg = ng;
#pragma unroll
        for (int j2 = 0; j2 < RA; j2++) cur[j2] = nxt[j2];
    }
}

// ---- direct circular convolution for short lines ------------------------------
// For n up to kDirectMax, an axis pass is the circular
// convolution out[i] = sum_j g[(i - j) mod n] in[j] with the exact periodic
// kernel g[m] = (1/n) sum_k f(k) cos(2 pi k m / n) of the spectral factor f
// (identical in exact arithmetic to the FFT -> multiply -> inverse FFT of the
// other path).  One thread owns one line: it loads the line into registers
// (FP64), forms the n outputs with n FMAs each (taps from the kernel
// arguments, i.e. scalar registers; g is even, so g[m] = g[n - m]) and
// stores them over the line.  No LDS, no barriers: the pass streams at
// memory speed, and its n FMAs per point match the FFT's flop count at
// these lengths.
constexpr int kDirectMax = 64;
constexpr int kDirectRowsMax = 56;  // x rows staged through LDS (rows kernel instances)
struct Taps {
    double g[kDirectMax / 2 + 1];
};
struct DirectPass {
    long long nlines;  // lines of one grid
    long long na;      // lines along the fast batch dimension
    long long sa, sb;  // start of line l: (l % na) * sa + (l / na) * sb
    long long step;    // stride along the line
    int zero_pads;     // x pass: zero columns [n, padw) of each row
    int padw;
};

// out[i * step] = sum_j g[(i - j) mod N] v[j], j ascending, for every i;
// eight outputs at a time so eight independent FMA chains interleave.
// The line is held in FP32 (as stored) and each value is widened to FP64
// once per block of outputs (exact), halving the registers of an FP64 copy.
template <int N>
__device__ inline void conv_line(const float (&v)[N], const Taps &taps, float *out,
                                 long long step) {
    constexpr int kB = 8;  // the last block of outputs may be partial
#pragma unroll
    for (int i0 = 0; i0 < N; i0 += kB) {
        double acc[kB];
#pragma unroll
        for (int b = 0; b < kB; b++) acc[b] = 0.0;
#pragma unroll
        for (int j = 0; j < N; j++) {
            const double vj = (double)v[j];
#pragma unroll
            for (int b = 0; b < kB; b++) {
                if (i0 + b < N) {
                    const int d = (i0 + b - j + N) % N;
                    acc[b] = fma(taps.g[d <= N / 2 ? d : N - d], vj, acc[b]);
                }
            }
        }
#pragma unroll
        for (int b = 0; b < kB; b++)
            if (i0 + b < N) out[(long long)(i0 + b) * step] = (float)acc[b];
    }
}

// Strided lines (y and z passes): consecutive threads own lines at
// consecutive x, so every load and store of the wave is coalesced.
template <int N>
__global__ void __launch_bounds__(256)
smooth_direct_k(float *grid, DirectPass dp, Taps taps, const long long *goffs,
                const GridLines *gl) {
    const long long line = (long long)blockIdx.x * 256 + threadIdx.x;
    float *g0 = grid + (goffs ? goffs[blockIdx.y] : 0LL);
    if (gl) {  // this grid's lines of a pass over several shapes
        const GridLines d = gl[blockIdx.y];
        g0 = grid + d.goff;
        dp.na = d.nb;
        dp.nlines = (long long)d.nb * d.nc;
        dp.sa = d.sb;
        dp.sb = d.sc;
        dp.step = d.sa;
    }
    if (line >= dp.nlines) return;
    float *p = g0 + (line % dp.na) * dp.sa + (line / dp.na) * dp.sb;
    float v[N];
#pragma unroll
    for (int j = 0; j < N; j++) v[j] = p[(long long)j * dp.step];
    conv_line<N>(v, taps, p, dp.step);
}

// Contiguous rows (x pass): the block's BLK consecutive rows (BLK * padw
// floats, contiguous in the grid) are staged through LDS with a row stride
// of padw + 1 (odd: the rows a wave reads start in different banks), so the
// global loads and stores are coalesced; pad columns are written as zero.
template <int N, int BLK>
__global__ void __launch_bounds__(BLK)
smooth_direct_x_k(float *grid, DirectPass dp, Taps taps, const long long *goffs,
                  const GridLines *gl) {
    extern __shared__ float tile[];
    // row width mx = 2 (N/2 + 1) (checked by the host)
    constexpr int W = 2 * (N / 2 + 1), WS = W + 1;
    const long long row0 = (long long)blockIdx.x * BLK;
    float *g0 = grid + (goffs ? goffs[blockIdx.y] : 0LL);
    if (gl) {  // this grid's rows of a pass over several shapes
        const GridLines d = gl[blockIdx.y];
        g0 = grid + d.goff;
        dp.nlines = (long long)d.nb * d.nc;
        if (row0 >= dp.nlines) return;  // the whole block
    }
    const int nrows = (int)(dp.nlines - row0 < BLK ? dp.nlines - row0 : BLK);
    float *base = g0 + row0 * W;
    const int tot = nrows * W;
    // kLB independent loads per thread in flight before their LDS stores
    constexpr int kLB = 12;
    for (int e0 = threadIdx.x; e0 < tot; e0 += BLK * kLB) {
        float v[kLB];
#pragma unroll
        for (int u = 0; u < kLB; u++) {
            const int e = e0 + u * BLK;
            v[u] = e < tot ? base[e] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kLB; u++) {
            const int e = e0 + u * BLK;
            if (e < tot) tile[(e / W) * WS + e % W] = v[u];
        }
    }
    __syncthreads();
    if (threadIdx.x < nrows) {
        float *r = tile + threadIdx.x * WS;
        float v[N];
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = r[j];
        conv_line<N>(v, taps, r, 1);
        for (int j = N; j < W; j++) r[j] = 0.0f;
    }
    __syncthreads();
    for (int e0 = threadIdx.x; e0 < tot; e0 += BLK * kLB) {
        float v[kLB];
#pragma unroll
        for (int u = 0; u < kLB; u++) {
            const int e = e0 + u * BLK;
            v[u] = e < tot ? tile[(e / W) * WS + e % W] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kLB; u++) {
            const int e = e0 + u * BLK;
            if (e < tot) base[e] = v[u];
        }
    }
}

// Whole cubic grids of side N <= 32 (N a multiple of 8), one workgroup per
// grid: the padded grid (N * N * (N + 2) floats, 139 KB at N = 32) is loaded
// into LDS once, the x, y and z passes run there, one line per thread, with
// the same conv_line as the per-pass kernels (so every value is the same
// bit for bit: the grid is FP32 between passes in both), and it is stored
// once -- a third of the HBM traffic of three passes.
template <int N>
__global__ void __launch_bounds__(1024) smooth_cube_k(float *grid, Taps taps,
                                                      const long long *goffs) {
    constexpr int W = N + 2;
    constexpr int CELLS = N * N * W;
    static_assert(CELLS % 2 == 0, "float pairs");
    __shared__ float c[CELLS];
    float *g = grid + (goffs ? goffs[blockIdx.x] : 0LL);
    // grid offsets and rows are even: 8-byte aligned float pairs
    const float2 *g2 = reinterpret_cast<const float2 *>(g);
    float2 *c2 = reinterpret_cast<float2 *>(c);
    for (int e = threadIdx.x; e < CELLS / 2; e += 1024) c2[e] = g2[e];
    __syncthreads();
    for (int r = threadIdx.x; r < N * N; r += 1024) {  // x: rows (y, z)
        float *p = c + r * W;
        float v[N];
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = p[j];
        conv_line<N>(v, taps, p, 1);
        p[N] = 0.0f;
        p[N + 1] = 0.0f;
    }
    __syncthreads();
    for (int l = threadIdx.x; l < N * N; l += 1024) {  // y: lines (x, z)
        float *p = c + l % N + (l / N) * (N * W);
        float v[N];
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = p[j * W];
        conv_line<N>(v, taps, p, W);
    }
    __syncthreads();
    for (int l = threadIdx.x; l < N * N; l += 1024) {  // z: lines (x, y)
        float *p = c + l % N + (l / N) * W;
        float v[N];
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = p[j * (N * W)];
        conv_line<N>(v, taps, p, N * W);
    }
    __syncthreads();
    float2 *o2 = reinterpret_cast<float2 *>(g);
    for (int e = threadIdx.x; e < CELLS / 2; e += 1024) o2[e] = c2[e];
}

// ---- host planning ----------------------------------------------------------
struct AxisPlan {
    int n = 0, L = 0;
    double2 *tw = nullptr;
    double *H = nullptr;
    bool direct = false;  // n <= kDirectMax: direct convolution possible
    Taps taps{};
};

std::mutex g_plan_mu;
std::map<std::tuple<int, int, int, unsigned long long>, AxisPlan> g_plans;

int next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// gsmooth.c:76-86: exp(k1*k1*expfac/(n*n)), k1 the signed frequency.
double axis_factor(long k, int n, double expfac) {
    long k1 = k > n / 2 ? k - n : k;
    return std::exp((double)(k1 * k1) * expfac / (double)((long)n * n));
}

hipError_t get_plan(int n, double expfac, AxisPlan &out) {
    int dev = 0;
    PGALF_HIP_TRY(hipGetDevice(&dev));
    unsigned long long eb;
    std::memcpy(&eb, &expfac, sizeof(eb));
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto key = std::make_tuple(dev, n, 0, eb);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
        out = it->second;
        return hipSuccess;
    }
    const bool pow2 = (n & (n - 1)) == 0;
    const int L = pow2 ? n : next_pow2(2 * n - 1);
    if (L > kMaxL) return hipErrorInvalidValue;
    const long double kPi = 3.141592653589793238462643383279502884L;
    std::vector<double2> tw(L);
    for (int m = 0; m < L; m++) {
        long double a = -2.0L * kPi * (long double)m / (long double)L;
        tw[m] = make_double2((double)cosl(a), (double)sinl(a));
    }
    std::vector<double> f(n);
    for (int k = 0; k < n; k++) f[k] = axis_factor(k, n, expfac);
    std::vector<double> H(L);
    // periodic kernel g[m] = (1/n) sum_k f(k) cos(2 pi k m / n)
    const bool direct = n <= kDirectMax;
    std::vector<long double> gk;
    if (direct || !pow2) {
        gk.resize(n);
        for (int m = 0; m < n; m++) {
            long double s = 0.0L;
            for (int k = 0; k < n; k++)
                s += (long double)f[k] *
                     cosl(2.0L * kPi * (long double)(((long)k * m) % n) / (long double)n);
            gk[m] = s / (long double)n;
        }
    }
    if (pow2) {
        for (int k = 0; k < n; k++) H[k] = f[k] / (double)n;
    } else {
        std::vector<long double> h(L, 0.0L);
        h[0] = gk[0];
        for (int t = 1; t < n; t++) {
            h[t] = gk[t];
            h[L - t] = gk[n - t];
        }
        for (int k = 0; k < L; k++) {
            long double s = 0.0L;
            for (int t = 0; t < L; t++)
                if (h[t] != 0.0L)
                    s += h[t] * cosl(2.0L * kPi * (long double)(((long)k * t) % L) /
                                     (long double)L);
            H[k] = (double)(s / (long double)L);
        }
    }
    AxisPlan p;
    p.n = n;
    p.L = L;
    p.direct = direct;
    if (direct)
        for (int m = 0; m <= n / 2; m++) p.taps.g[m] = (double)gk[m];
    PGALF_HIP_TRY(hipMalloc(&p.tw, sizeof(double2) * L));
    PGALF_HIP_TRY(hipMalloc(&p.H, sizeof(double) * L));
    PGALF_HIP_TRY(hipMemcpy(p.tw, tw.data(), sizeof(double2) * L, hipMemcpyHostToDevice));
    PGALF_HIP_TRY(hipMemcpy(p.H, H.data(), sizeof(double) * L, hipMemcpyHostToDevice));
    g_plans[key] = p;
    out = p;
    return hipSuccess;
}

// Kernel choice per line length: the direct convolution for lengths up to
// the schedule's smooth_direct_max (rows: up to kDirectRowsMax as well),
// else the register FFT for n = 64, 128, 256, else the Stockham kernel.
// The default 56 takes 1..56 directly and 64 by the register FFT, which
// measured faster in the batch.
#ifndef PGALF_SMOOTH_FFT2
#define PGALF_SMOOTH_FFT2 1
#endif
#ifndef PGALF_SMOOTH_CUBE
#define PGALF_SMOOTH_CUBE 1  // small cubes in one LDS-resident kernel
#endif
#ifndef PGALF_FFT2_NL256
#define PGALF_FFT2_NL256 16  // complex lines per block at L = 256 (16: 128-byte rows)
#endif

// Resident blocks of one kernel instance on the current device (cached).
template <class K>
int resident_blocks(K kernel, int threads) {
    static std::mutex mu;
    static std::map<int, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess ||
        cus < 1 || per < 1)
        return 1024;
    cache[dev] = cus * per;
    return cus * per;
}

hipError_t run_fft2(float *g, const AxisPlan &pl, const AxisPass &ap, int nc, hipStream_t s,
                    const long long *goffs, int ngrids) {
#define PGALF_FFT2(RA, RB, NL)                                                                 \
    {                                                                                          \
        Fft2Groups gr;                                                                         \
        gr.nbx = (ap.nb + 2 * (NL) - 1) / (2 * (NL));                                          \
        gr.nc = nc;                                                                            \
        const long long ng = (long long)gr.nbx * nc * ngrids;                                 \
        if (ng >= (1LL << 31)) return hipErrorInvalidValue;                                    \
        gr.ngroups = (unsigned)ng;                                                             \
        constexpr int kThreads = (NL) * (RB);                                                  \
        auto kern = ap.contiguous ? smooth_fft2_k<RA, RB, NL, true>                            \
                                  : smooth_fft2_k<RA, RB, NL, false>;                          \
        long long blocks = resident_blocks(kern, kThreads);                                    \
        if (blocks > (long long)gr.ngroups) blocks = gr.ngroups;                               \
        kern<<<(unsigned)blocks, kThreads, 0, s>>>(g, ap, pl.tw, pl.H, goffs, gr);             \
    }
    switch (pl.L) {
    case 64: PGALF_FFT2(8, 8, 32) break;
    case 128: PGALF_FFT2(8, 16, 16) break;
    case 256: PGALF_FFT2(16, 16, PGALF_FFT2_NL256) break;
    default: return hipErrorInvalidValue;
    }
#undef PGALF_FFT2
    return hipGetLastError();
}

// A pass over several grids of different shapes (a batch's grids with this
// line length): per-grid lines in gl[0 .. count), the launch sized by the
// largest grid; sa == 1 marks the x pass, nb is the smallest grid's line
// count (it caps the lines per block).
struct MixedLines {
    const GridLines *gl = nullptr;
    int count = 0;
    int min_nb = 0, max_nb = 0, max_nc = 0;
    long long max_lines = 0;
};

bool fft2_length(const AxisPlan &pl, long long gcells) {
    return PGALF_SMOOTH_FFT2 && pl.L == pl.n && (pl.n == 64 || pl.n == 128 || pl.n == 256) &&
           gcells > 0 && gcells < (1LL << 29);
}
bool direct_length(const AxisPlan &pl, bool rows, int dmax) {
    return pl.direct && pl.n <= dmax && (!rows || pl.n <= kDirectRowsMax);
}

hipError_t run_pass(float *g, const AxisPlan &pl, long long sa, int nb, long long sb, int nc,
                    long long sc, int contiguous, int zero_pads, int padw, long long gcells,
                    hipStream_t s, const long long *goffs = nullptr, int ngrids = 1,
                    const MixedLines *mix = nullptr, int dmax = -1) {
    if (dmax < 0) dmax = smooth_direct_max();
    const bool fft2 = !mix && fft2_length(pl, gcells);
    const bool direct = direct_length(pl, sa == 1, dmax);
    if (mix) ngrids = mix->count;
    if (fft2 && !direct) {
        AxisPass ap{};
        ap.n = pl.n;
        ap.L = pl.L;
        ap.sa = sa;
        ap.nb = nb;
        ap.sb = sb;
        ap.sc = sc;
        ap.contiguous = contiguous;
        ap.zero_pads = zero_pads;
        ap.padw = padw;
        // grid offsets and slow strides are even (mx is), so line pairs
        // starting at even x are 8-byte aligned when the grid is
        ap.pairs = !contiguous && sb == 1 && (sc % 2 == 0) &&
                   ((reinterpret_cast<unsigned long long>(g) & 7ull) == 0);
        ap.gcells = gcells;
        return run_fft2(g, pl, ap, nc, s, goffs, ngrids);
    }
    if (direct) {
        DirectPass dp;
        dp.na = nb;
        dp.nlines = (long long)nb * nc;
        dp.sa = sb;
        dp.sb = sc;
        dp.step = sa;
        dp.zero_pads = zero_pads;
        dp.padw = padw;
        constexpr int kMaxY = 65535;
        // x pass: rows staged through LDS, 256 rows per block
        const bool rows = sa == 1;
        if (rows && padw != 2 * (pl.n / 2 + 1)) return hipErrorInvalidValue;
        const int blk = 256;
        const long long lines = mix ? mix->max_lines : dp.nlines;
        const unsigned gx = (unsigned)((lines + blk - 1) / blk);
        const size_t lds = rows ? sizeof(float) * (size_t)blk * (padw + 1) : 0;
        for (int z0 = 0; z0 < ngrids; z0 += kMaxY) {
            const int nz = ngrids - z0 < kMaxY ? ngrids - z0 : kMaxY;
            const long long *gz = goffs && !mix ? goffs + z0 : nullptr;
            const GridLines *gl = mix ? mix->gl + z0 : nullptr;
            const dim3 grid(gx, (unsigned)nz);
#define PGALF_DIRECT_ROWS(NN)                                                                  \
    case NN:                                                                                   \
        if (rows)                                                                              \
            smooth_direct_x_k<NN, 256><<<grid, 256, lds, s>>>(g, dp, pl.taps, gz, gl);         \
        else                                                                                   \
            smooth_direct_k<NN><<<grid, 256, 0, s>>>(g, dp, pl.taps, gz, gl);                  \
        break;
#define PGALF_DIRECT_STRIDED(NN)                                                               \
    case NN:                                                                                   \
        smooth_direct_k<NN><<<grid, 256, 0, s>>>(g, dp, pl.taps, gz, gl);                      \
        break;
            switch (pl.n) {
                PGALF_DIRECT_ROWS(1) PGALF_DIRECT_ROWS(2) PGALF_DIRECT_ROWS(3)
                PGALF_DIRECT_ROWS(4) PGALF_DIRECT_ROWS(5) PGALF_DIRECT_ROWS(6)
                PGALF_DIRECT_ROWS(7) PGALF_DIRECT_ROWS(8) PGALF_DIRECT_ROWS(9)
                PGALF_DIRECT_ROWS(10) PGALF_DIRECT_ROWS(11) PGALF_DIRECT_ROWS(12)
                PGALF_DIRECT_ROWS(13) PGALF_DIRECT_ROWS(14) PGALF_DIRECT_ROWS(15)
                PGALF_DIRECT_ROWS(16) PGALF_DIRECT_ROWS(17) PGALF_DIRECT_ROWS(18)
                PGALF_DIRECT_ROWS(19) PGALF_DIRECT_ROWS(20) PGALF_DIRECT_ROWS(21)
                PGALF_DIRECT_ROWS(22) PGALF_DIRECT_ROWS(23) PGALF_DIRECT_ROWS(24)
                PGALF_DIRECT_ROWS(25) PGALF_DIRECT_ROWS(26) PGALF_DIRECT_ROWS(27)
                PGALF_DIRECT_ROWS(28) PGALF_DIRECT_ROWS(29) PGALF_DIRECT_ROWS(30)
                PGALF_DIRECT_ROWS(31) PGALF_DIRECT_ROWS(32) PGALF_DIRECT_ROWS(33)
                PGALF_DIRECT_ROWS(34) PGALF_DIRECT_ROWS(35) PGALF_DIRECT_ROWS(36)
                PGALF_DIRECT_ROWS(37) PGALF_DIRECT_ROWS(38) PGALF_DIRECT_ROWS(39)
                PGALF_DIRECT_ROWS(40) PGALF_DIRECT_ROWS(41) PGALF_DIRECT_ROWS(42)
                PGALF_DIRECT_ROWS(43) PGALF_DIRECT_ROWS(44) PGALF_DIRECT_ROWS(45)
                PGALF_DIRECT_ROWS(46) PGALF_DIRECT_ROWS(47) PGALF_DIRECT_ROWS(48)
                PGALF_DIRECT_ROWS(49) PGALF_DIRECT_ROWS(50) PGALF_DIRECT_ROWS(51)
                PGALF_DIRECT_ROWS(52) PGALF_DIRECT_ROWS(53) PGALF_DIRECT_ROWS(54)
                PGALF_DIRECT_ROWS(55) PGALF_DIRECT_ROWS(56)
                PGALF_DIRECT_STRIDED(57) PGALF_DIRECT_STRIDED(58) PGALF_DIRECT_STRIDED(59)
                PGALF_DIRECT_STRIDED(60) PGALF_DIRECT_STRIDED(61) PGALF_DIRECT_STRIDED(62)
                PGALF_DIRECT_STRIDED(63) PGALF_DIRECT_STRIDED(64)
            default: return hipErrorInvalidValue;
            }
#undef PGALF_DIRECT_ROWS
#undef PGALF_DIRECT_STRIDED
        }
        return hipGetLastError();
    }
    AxisPass ap{};
    ap.n = pl.n;
    ap.L = pl.L;
    ap.sa = sa;
    ap.nb = nb;
    ap.sb = sb;
    ap.sc = sc;
    // P * L <= LC bounds the butterflies per thread of a pass
    const int LC = pl.L > kLdsComplex ? kLdsComplexMax : kLdsComplex;
    ap.tables = pl.L <= kLdsTables ? 1 : 0;
    const size_t tbytes = ap.tables ? (sizeof(double2) + sizeof(double)) * (size_t)pl.L : 0;
    ap.LS = pl.L + 1;
    int P = (int)((kLdsBytes - tbytes) / (sizeof(double2) * (size_t)ap.LS));
    if (P > LC / pl.L) P = LC / pl.L;
    if (P < 1) {  // one line of the longest length fills the block unpadded
        P = 1;
        ap.LS = pl.L;
    }
    int p2 = 1;  // a power of two: line counts (grid sides) divide evenly
    while (2 * p2 <= P) p2 *= 2;
    P = p2;
    int maxP = (nb + 1) / 2;
    if (P > maxP) P = maxP;
    ap.P = P;
    ap.contiguous = contiguous;
    ap.zero_pads = zero_pads;
    ap.padw = padw;
    const size_t lds = sizeof(double2) * (size_t)P * ap.LS + tbytes;
    constexpr int kMaxZ = 65535;  // grid dimension limit: larger batches in several launches
    const int gnb = mix ? mix->max_nb : nb, gnc = mix ? mix->max_nc : nc;
    for (int z0 = 0; z0 < ngrids; z0 += kMaxZ) {
        const int nz = ngrids - z0 < kMaxZ ? ngrids - z0 : kMaxZ;
        dim3 grid((unsigned)((gnb + 2 * P - 1) / (2 * P)), (unsigned)gnc, (unsigned)nz);
        const long long *gz = goffs && !mix ? goffs + z0 : nullptr;
        const GridLines *gl = mix ? mix->gl + z0 : nullptr;
        if (LC == kLdsComplex)
            smooth_axis_k<kLdsComplex><<<grid, kBlock, lds, s>>>(g, ap, pl.tw, pl.H, gz, gl);
        else
            smooth_axis_k<kLdsComplexMax><<<grid, kBlock, lds, s>>>(g, ap, pl.tw, pl.H, gz, gl);
    }
    return hipGetLastError();
}

// The axis pass `axis` (0: x rows, 1: y, 2: z) of grids of shape nx, ny, nz
// at goffs (or, with mix, of the grids in mix->gl, of any shape with this
// line length; nx, ny, nz then only give the line length and row width).
hipError_t axis_pass(float *den, int axis, const AxisPlan &pl, int nx, int ny, int nz,
                     hipStream_t s, const long long *goffs, int ngrids,
                     const MixedLines *mix = nullptr, int dmax = -1) {
    const long long mx = 2LL * (nx / 2 + 1);
    const long long gc = mx * ny * nz;
    switch (axis) {
    case 0:  // lines are rows (contiguous), batch over ny*nz rows
        return run_pass(den, pl, 1, mix ? mix->min_nb : ny * nz, mx, 1, 0, 1, 1, (int)mx, gc,
                        s, goffs, ngrids, mix, dmax);
    case 1:  // lines along y (stride mx), batch over x (stride 1), one z per block row
        return run_pass(den, pl, mx, mix ? mix->min_nb : nx, 1, nz, mx * ny, 0, 0, 0, gc, s,
                        goffs, ngrids, mix, dmax);
    default:  // lines along z (stride mx*ny), batch over x, one y per block row
        return run_pass(den, pl, mx * ny, mix ? mix->min_nb : nx, 1, ny, mx, 0, 0, 0, gc, s,
                        goffs, ngrids, mix, dmax);
    }
}

// Small cubes whose three passes are all direct: one LDS-resident kernel.
bool cube_shape(int nx, int ny, int nz, const AxisPlan &px, int dmax) {
    return PGALF_SMOOTH_CUBE && nx == ny && ny == nz && nx % 8 == 0 && px.direct && nx <= 32 &&
           nx <= dmax;
}
hipError_t cube_pass(float *den, int n, const Taps &taps, hipStream_t s, const long long *goffs,
                     int ngrids) {
    for (int z0 = 0; z0 < ngrids; z0 += 1 << 30) {
        const int cnt = ngrids - z0 < (1 << 30) ? ngrids - z0 : (1 << 30);
        const long long *gz = goffs ? goffs + z0 : nullptr;
        switch (n) {
        case 8: smooth_cube_k<8><<<cnt, 1024, 0, s>>>(den, taps, gz); break;
        case 16: smooth_cube_k<16><<<cnt, 1024, 0, s>>>(den, taps, gz); break;
        case 24: smooth_cube_k<24><<<cnt, 1024, 0, s>>>(den, taps, gz); break;
        case 32: smooth_cube_k<32><<<cnt, 1024, 0, s>>>(den, taps, gz); break;
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

// The three axis passes of one grid shape over `ngrids` grids.
hipError_t smooth_passes(float *den, int nx, int ny, int nz, double expfac, hipStream_t s,
                         const long long *goffs, int ngrids) {
    AxisPlan p[3];
    PGALF_HIP_TRY(get_plan(nx, expfac, p[0]));
    PGALF_HIP_TRY(get_plan(ny, expfac, p[1]));
    PGALF_HIP_TRY(get_plan(nz, expfac, p[2]));
    const int dmax = smooth_direct_max();
    if (cube_shape(nx, ny, nz, p[0], dmax)) return cube_pass(den, nx, p[0].taps, s, goffs, ngrids);
    for (int a = 0; a < 3; a++)
        PGALF_HIP_TRY(axis_pass(den, a, p[a], nx, ny, nz, s, goffs, ngrids, nullptr, dmax));
    return hipSuccess;
}

}  // namespace

double smooth_expfac(float RG, double cellsize) {
    // gsmooth.c:67-69: RG rescaled to cells in float, expfac in double.
    float rgc = (float)((double)RG / cellsize);
    return -(M_PI * (double)rgc) * (M_PI * (double)rgc);
}

hipError_t gaussian_smooth(float *den, const Grid &g, float RG, Workspace &ws, hipStream_t s) {
    (void)ws;
    return smooth_passes(den, g.nx, g.ny, g.nz, smooth_expfac(RG, g.h), s, nullptr, 1);
}

hipError_t smooth_group_prepare(const SmoothGroup &grp) {
    AxisPlan p;
    PGALF_HIP_TRY(get_plan(grp.nx, grp.expfac, p));
    PGALF_HIP_TRY(get_plan(grp.ny, grp.expfac, p));
    return get_plan(grp.nz, grp.expfac, p);
}

// A catalogue's smoothing: small cubes by the LDS-resident kernel per shape;
// the axis passes of lengths that take the register FFT per shape (its
// launch geometry is per shape); every other axis pass -- direct
// convolution or the Stockham FFT -- as one launch per (axis, line length,
// smoothing factor) over the grids of all shapes that have it, so that a
// catalogue of many shapes does not pay one small launch per shape and
// axis.  Passes run axis by axis (all x, then all y, then all z), which
// keeps each grid's x -> y -> z order.  The direct-convolution threshold is
// the one in force when the plan is made.
struct SmoothBatchPlan {
    int dmax = 56;
    std::vector<int> cubes;      // groups
    std::vector<int> whole[3];   // groups whose axis pass runs per shape
    struct Mixed {
        int axis, nx, ny, nz;  // a representative shape (line length, row width)
        double expfac;
        size_t first;
        MixedLines ml;
    };
    std::vector<Mixed> mixed;
    GridLines *d_lines = nullptr;
};

SmoothBatchPlan *smooth_batch_plan_create(const SmoothGroup *groups, int ngroups,
                                          hipError_t &err) {
    err = hipSuccess;
    auto *pl = new SmoothBatchPlan;
    pl->dmax = smooth_direct_max();
    std::map<std::tuple<int, int, unsigned long long>, std::vector<GridLines>> mix;
    std::map<std::tuple<int, int, unsigned long long>, std::array<int, 3>> rep;
    for (int i = 0; i < ngroups; i++) {
        const SmoothGroup &g = groups[i];
        AxisPlan p[3];
        const int dims[3] = {g.nx, g.ny, g.nz};
        for (int a = 0; a < 3 && err == hipSuccess; a++) err = get_plan(dims[a], g.expfac, p[a]);
        if (err != hipSuccess) break;
        if (cube_shape(g.nx, g.ny, g.nz, p[0], pl->dmax)) {
            pl->cubes.push_back(i);
            continue;
        }
        const long long mx = 2LL * (g.nx / 2 + 1), gc = mx * g.ny * g.nz;
        unsigned long long eb;
        std::memcpy(&eb, &g.expfac, sizeof(eb));
        for (int a = 0; a < 3; a++) {
            if (fft2_length(p[a], gc) && !direct_length(p[a], a == 0, pl->dmax)) {
                pl->whole[a].push_back(i);
                continue;
            }
            const auto key = std::make_tuple(a, dims[a], eb);
            rep[key] = {g.nx, g.ny, g.nz};
            auto &v = mix[key];
            for (int k = 0; k < g.count; k++) {
                GridLines d;
                d.goff = g.hgoffs[k];
                if (a == 0) {
                    d.sa = 1; d.sb = mx; d.sc = 0; d.nb = g.ny * g.nz; d.nc = 1;
                } else if (a == 1) {
                    d.sa = mx; d.sb = 1; d.sc = mx * g.ny; d.nb = g.nx; d.nc = g.nz;
                } else {
                    d.sa = mx * g.ny; d.sb = 1; d.sc = mx; d.nb = g.nx; d.nc = g.ny;
                }
                v.push_back(d);
            }
        }
    }
    std::vector<GridLines> all;
    for (auto &kv : mix) {
        SmoothBatchPlan::Mixed m;
        m.axis = std::get<0>(kv.first);
        const auto &r = rep[kv.first];
        m.nx = r[0];
        m.ny = r[1];
        m.nz = r[2];
        std::memcpy(&m.expfac, &std::get<2>(kv.first), sizeof(m.expfac));
        m.first = all.size();
        m.ml.count = (int)kv.second.size();
        m.ml.min_nb = 1 << 30;
        for (const GridLines &d : kv.second) {
            m.ml.min_nb = std::min(m.ml.min_nb, d.nb);
            m.ml.max_nb = std::max(m.ml.max_nb, d.nb);
            m.ml.max_nc = std::max(m.ml.max_nc, d.nc);
            m.ml.max_lines = std::max(m.ml.max_lines, (long long)d.nb * d.nc);
        }
        all.insert(all.end(), kv.second.begin(), kv.second.end());
        pl->mixed.push_back(m);
    }
    if (err == hipSuccess && !all.empty()) {
        err = hipMalloc(&pl->d_lines, sizeof(GridLines) * all.size());
        if (err == hipSuccess)
            err = hipMemcpy(pl->d_lines, all.data(), sizeof(GridLines) * all.size(),
                            hipMemcpyHostToDevice);
    }
    for (auto &m : pl->mixed) m.ml.gl = pl->d_lines + m.first;
    if (err != hipSuccess) {
        smooth_batch_plan_free(pl);
        return nullptr;
    }
    return pl;
}

void smooth_batch_plan_free(SmoothBatchPlan *pl) {
    if (!pl) return;
    if (pl->d_lines) (void)hipFree(pl->d_lines);
    delete pl;
}

hipError_t gaussian_smooth_batch(float *den, const SmoothGroup *groups, int ngroups,
                                 const SmoothBatchPlan *pl, hipStream_t s) {
    (void)ngroups;
    for (int i : pl->cubes) {
        const SmoothGroup &g = groups[i];
        AxisPlan p;
        PGALF_HIP_TRY(get_plan(g.nx, g.expfac, p));
        PGALF_HIP_TRY(cube_pass(den, g.nx, p.taps, s, g.goffs, g.count));
    }
    for (int a = 0; a < 3; a++) {
        for (int i : pl->whole[a]) {
            const SmoothGroup &g = groups[i];
            const int dims[3] = {g.nx, g.ny, g.nz};
            AxisPlan p;
            PGALF_HIP_TRY(get_plan(dims[a], g.expfac, p));
            PGALF_HIP_TRY(axis_pass(den, a, p, g.nx, g.ny, g.nz, s, g.goffs, g.count, nullptr,
                                    pl->dmax));
        }
        for (const auto &m : pl->mixed) {
            if (m.axis != a) continue;
            const int dims[3] = {m.nx, m.ny, m.nz};
            AxisPlan p;
            PGALF_HIP_TRY(get_plan(dims[a], m.expfac, p));
            PGALF_HIP_TRY(axis_pass(den, a, p, m.nx, m.ny, m.nz, s, nullptr, m.ml.count, &m.ml,
                                    pl->dmax));
        }
    }
    return hipSuccess;
}

}  // namespace pgalf
