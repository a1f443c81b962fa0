// Internal declarations of the MI355X (gfx950) density-stage library.
// Kernels live in tsc.hip, smooth.hip, knn.hip and scan_sort.hip; the C-ABI
// in capi.hip.  All device memory beyond caller buffers comes from one
// per-device Workspace arena.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/pgalf_density.h"

namespace pgalf {

// Grid geometry of the reference's padded in-place r2c layout:
// den(i,j,k) = den[i + mx*(j + ny*k)], mx = 2*(nx/2+1) (tsc_omp2.c:16,198).
struct Grid {
    int nx, ny, nz;
    double x0, y0, z0, h;
    __host__ __device__ long long mx() const { return 2LL * (nx / 2 + 1); }
    __host__ __device__ long long ncells() const { return mx() * (long long)ny * nz; }
};

// Device error bits raised by kernels (read back with the status word).
enum : unsigned {
    DEV_ERR_OUT_OF_GRID = 1u,
    DEV_ERR_SHORT_LIST = 2u,
};

// Grow-only device arena.  begin() resets the bump pointer; alloc() never
// moves earlier blocks (extra chunks are added and coalesced at the next
// begin(), after the stream has drained).
class Workspace {
  public:
    explicit Workspace(int device) : device_(device) {}
    ~Workspace();
    hipError_t begin(hipStream_t s);
    void *alloc(size_t bytes);
    template <class T> T *alloc_n(size_t n) { return static_cast<T *>(alloc(n * sizeof(T))); }
    unsigned *status() { return status_; }  // device word of DEV_ERR_* bits
    int device() const { return device_; }

  private:
    struct Chunk { char *ptr; size_t cap, used; };
    std::vector<Chunk> chunks_;
    size_t high_water_ = 0;
    unsigned *status_ = nullptr;
    int device_;
    hipError_t add_chunk(size_t bytes);
};

Workspace &workspace_for_current_device();

// ---- scan / sort (scan_sort.hip) ----------------------------------------
// Exclusive scan of n uint32 values; out may alias in.  If total != nullptr
// the sum is written there (device pointer).
hipError_t exclusive_scan_u32(const unsigned *in, unsigned *out, long long n,
                              unsigned *total, Workspace &ws, hipStream_t s);
// Stable LSD radix sort of (key, value) pairs on the low `bits` key bits.
// Sorted data is returned in *keys/*vals (the pointers may be swapped with
// the tmp buffers).
hipError_t radix_sort_pairs(unsigned **keys, unsigned **vals, unsigned *keys_tmp,
                            unsigned *vals_tmp, long long n, int bits,
                            Workspace &ws, hipStream_t s);

// ---- TSC deposit / read-back (tsc.hip) ----------------------------------
template <class P>
hipError_t tsc_deposit(const P *x, const P *y, const P *z, const P *m,
                       const int *type, long long np, int ptype, const Grid &g,
                       float *den, Workspace &ws, hipStream_t s);
template <class P>
hipError_t tsc_readback(const P *x, const P *y, const P *z, long long np,
                        const Grid &g, const float *den, float *out,
                        Workspace &ws, hipStream_t s);

// ---- Gaussian smoothing (smooth.hip) ------------------------------------
hipError_t gaussian_smooth(float *den, const Grid &g, float RG, Workspace &ws,
                           hipStream_t s);

// ---- kNN (knn.hip) -------------------------------------------------------
// Exact k nearest candidates of every query (self included when the query
// set is the candidate set), sorted by (FP64 dist2, candidate index).
// nbr (int64 [nq*k], candidate indices) and rho (float [nq]) are optional.
// rho_mode: 0 none, 1 unweighted W4 sum, 2 W4 weighted by candidate mass.
// cand_sel: optional int type array + ptype selecting the candidates among
// the same particle arrays (nullptr: every particle is a candidate).
template <class P>
hipError_t knn(const P *x, const P *y, const P *z, const P *m, long long np,
               const int *type, int cand_ptype, int k, long long *nbr,
               float *rho, int rho_mode, Workspace &ws, hipStream_t s);

// AoS SimpleBasicParticleType (88 B) -> SoA double arrays on device.
hipError_t unpack_particles(const void *d_bp, long long np, double *x, double *y,
                            double *z, double *m, int *type, hipStream_t s);

// ---- phase timing (capi.hip) ---------------------------------------------
void timing_mark(const char *phase, hipStream_t s, bool end);

}  // namespace pgalf

#define PGALF_HIP_TRY(expr)                       \
    do {                                          \
        hipError_t _e = (expr);                   \
        if (_e != hipSuccess) return _e;          \
    } while (0)
