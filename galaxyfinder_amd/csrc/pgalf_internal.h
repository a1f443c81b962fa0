// Internal declarations of the MI355X (gfx950) density-stage library.
// Kernels live in tsc.hip, smooth.hip, knn.hip and scan_sort.hip; the C-ABI
// in capi.hip.  All device memory beyond caller buffers comes from one
// per-device Workspace arena.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/pgalf_density.h"

namespace pgalf {

// Grids are read and written as float pairs (mx and every grid offset are
// even): device grid pointers must be 8-byte aligned.
inline bool aligned8(const void *p) { return ((uintptr_t)p & 7u) == 0; }

// Grid geometry of the reference's padded in-place r2c layout:
// den(i,j,k) = den[i + mx*(j + ny*k)], mx = 2*(nx/2+1) (tsc_omp2.c:16,198).
struct Grid {
    int nx, ny, nz;
    double x0, y0, z0, h;
    __host__ __device__ long long mx() const { return 2LL * (nx / 2 + 1); }
    __host__ __device__ long long ncells() const { return mx() * (long long)ny * nz; }
};

// One halo of a batch (a catalogue processed in one pass).  Its particles
// are [p0, p0 + np) of the concatenated SoA arrays, its padded grid starts at
// float goff of the batch grid buffer and its deposit tiles at toff of the
// batch tile space.  The single-halo entry points use one HaloSeg with
// p0 = goff = toff = 0.
struct HaloSeg {
    long long p0, np;
    long long goff;
    long long toff;
    Grid g;
};

// A block of at most kChunkParticles particles that all belong to one halo.
struct PChunk {
    int halo;
    long long b0, b1;
};
constexpr int kChunkParticles = 8192;
// Finer chunks (one workgroup each) for the per-particle gathers of a batch:
// most halos of a catalogue are far smaller than kChunkParticles.
constexpr int kFineParticles = 256;

// Launch schedule of a workload (knobs that an autotuner sets; none of them
// changes a result bit except smooth_direct_max, whose methods agree within
// the smoothing tolerance).  schedule_for(n): the first bin of the active
// table with n <= nmax, the defaults beyond the last bin or with no table.
struct Schedule {
    long long nmax = -1;
    int knn_packed = 1;           // packed-key query (+ exact re-run of flagged waves)
    int knn_xcd_chunk = 64;       // consecutive query waves per XCD run (0: no remap)
    int tsc_unit_entries = 4096;  // bin entries per deposit work unit
};
Schedule schedule_for(long long n);
// Strided smoothing lines of length n (a multiple of 8) up to this use the
// direct convolution; rows up to min(this, 32).  Global, not binned.
int smooth_direct_max();

// Device view of a batch: descriptors built on the host once per catalogue.
struct BatchDev {
    Schedule sched;                 // fixed when the batch is created
    int nhalo = 0;
    long long np = 0;               // particles in all halos
    const HaloSeg *segs = nullptr;  // [nhalo]
    const PChunk *chunks = nullptr; // [nchunks]
    long long nchunks = 0;
    const PChunk *fine = nullptr;   // [nfine] chunks of <= kFineParticles
    long long nfine = 0;
    const unsigned *tile_halo = nullptr;  // [ntiles] halo of each deposit tile
    long long ntiles = 0;
    long long max_halo_tiles = 0;
    int max_tile_entries = 4;       // deposit tiles a particle can touch, at most
};

// Device error bits raised by kernels (read back with the status word).
enum : unsigned {
    DEV_ERR_OUT_OF_GRID = 1u,
    DEV_ERR_SHORT_LIST = 2u,
    DEV_ERR_BAD_INDEX = 4u,  // an index argument outside the particle range
};

// Grow-only device arena.  begin() resets the bump pointer; alloc() never
// moves earlier blocks (extra chunks are added and coalesced at the next
// begin(), after the stream has drained).
//
// Calls may come on different streams.  Every call ends with end(s), which
// records an event on its stream; a call on another stream first makes its
// stream wait for that event, so its kernels reuse the arena only after the
// previous call's kernels are done with it (calls on one stream are ordered
// by the stream itself).
class Workspace {
  public:
    // A bump position; release() returns the arena to it, so a stage's
    // scratch is reused by the stages after it.  Within one call every user
    // runs on one stream (or one captured graph), so later kernels that reuse
    // the memory run after the earlier ones have finished with it.
    struct Mark {
        size_t chunk, used;
    };
    explicit Workspace(int device) : device_(device) {}
    // A fixed arena over caller-owned device memory [base, base + bytes)
    // (256-byte aligned): it never allocates, frees or creates events; the
    // status word is its first 256 bytes, and an allocation that does not
    // fit returns nullptr.
    Workspace(int device, void *base, size_t bytes);
    ~Workspace();
    bool fixed() const { return fixed_; }
    hipError_t begin(hipStream_t s);
    // The call on stream s has queued all its work (see the class comment).
    void end(hipStream_t s);
    void *alloc(size_t bytes);
    template <class T> T *alloc_n(size_t n) { return static_cast<T *>(alloc(n * sizeof(T))); }
    Mark mark() const { return chunks_.empty() ? Mark{0, 0} : Mark{cur_, chunks_[cur_].used}; }
    void release(const Mark &m);
    unsigned *status() { return status_; }  // device word of DEV_ERR_* bits
    // Kernels run with this workspace report into another workspace's status
    // word (a side stream's scratch beside that workspace's own call).
    void borrow_status(unsigned *st) {
        if (own_status_ && status_) return;  // already has a word of its own
        status_ = st;
        own_status_ = false;
    }
    int device() const { return device_; }
    // Largest live scratch of any call so far (bytes), and the arena's size.
    size_t high_water() const { return peak_ > high_water_ ? peak_ : high_water_; }
    size_t reserved() const {
        size_t t = 0;
        for (const auto &c : chunks_) t += c.cap;
        return t;
    }

  private:
    struct Chunk { char *ptr; size_t cap, used; };
    std::vector<Chunk> chunks_;
    size_t cur_ = 0;         // chunk of the bump position
    size_t peak_ = 0;        // largest live total since begin()
    size_t high_water_ = 0;  // largest peak of any call
    unsigned *status_ = nullptr;
    int device_;
    hipEvent_t done_ = nullptr;  // recorded by end(): the last call's work
    hipStream_t last_ = nullptr;
    bool has_last_ = false;
    bool fixed_ = false;
    bool own_status_ = true;
    hipError_t add_chunk(size_t bytes);
};

// begin() on construction, end() when the entry point returns.
struct WsCall {
    Workspace &ws;
    hipStream_t s;
    hipError_t e;
    WsCall(Workspace &w, hipStream_t st) : ws(w), s(st), e(w.begin(st)) {}
    ~WsCall() { ws.end(s); }
    WsCall(const WsCall &) = delete;
    WsCall &operator=(const WsCall &) = delete;
};

// Releases a stage's workspace scratch when the stage returns (its kernels
// are queued; later stages on the same stream reuse the memory after them).
struct WsScope {
    Workspace &ws;
    Workspace::Mark m;
    explicit WsScope(Workspace &w) : ws(w), m(w.mark()) {}
    ~WsScope() { ws.release(m); }
    WsScope(const WsScope &) = delete;
    WsScope &operator=(const WsScope &) = delete;
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
// Fill n words with v (a kernel, not hipMemsetAsync: graph-capture safe).
hipError_t fill_u32(unsigned *p, long long n, unsigned v, hipStream_t s);
hipError_t fill_u64(unsigned long long *p, long long n, unsigned long long v, hipStream_t s);

// ---- watershed of subhalo_den (watershed.hip) ---------------------------
// FindCoreDensity outputs (device arrays): per core, per particle.
struct CoreOut {
    float *coredensity;
    int *nummem, *numstar;
    float *starmass;
    int *haloid;
    unsigned char *flag;
};
// Shells as CSR: particles part[part_ptr[s] .. part_ptr[s+1]), cores
// cores[core_ptr[s] .. core_ptr[s+1]); at most maxshell shells, maxcores
// core entries.
struct ShellOut {
    long long *part_ptr;
    int *part;
    long long *core_ptr;
    int *cores;
    long long maxshell, maxcores;
};
// Host-synchronising (reads counts back between stages); not capturable.
// *search_fail: a threshold search that never converged (the reference loops
// for ever there).
hipError_t find_core_density(const float *den, long long np, const long long *nbr, int k,
                             const int *type, const double *mass, const int *peak, int numcore,
                             const CoreOut &o, int *search_fail, Workspace &ws, hipStream_t s);
// The shell thresholds (host array thr[15], *nthr of them).
hipError_t shell_thresholds(const float *den, long long np, const float *coredensity,
                            int numcore, float *thr, int *nthr, Workspace &ws, hipStream_t s);
// Catalogue watershed: halo h's particles [off[h], off[h+1]) and cores
// [coff[h], coff[h+1]) (host arrays), lists and core peaks local to the
// halo.  Shells of halo h: part_ptr / core_ptr entries at sbase[h] (nc*nsd +
// 2 of them), its particles at off[h] of part, core entries at cbase[h]
// (ccap[h] of them) -- the layout catalog_shell_layout gives.
struct CatShellOut {
    long long *part_ptr;
    int *part;
    long long *core_ptr;
    int *cores;
};
void catalog_shell_layout(int nhalo, const long long *off, const long long *coff,
                          long long *sbase, long long *cbase, long long *ccap,
                          long long *tot_shell, long long *tot_cores);
// Halos of at most small_max particles run one per workgroup; the others
// (and any that outgrow it) through the one-halo tiers.  herr_h[h]: 0, or a
// non-zero code for a halo whose results are not valid (a search that did
// not converge: 2; shells: 10 + the core_shells error).
hipError_t catalog_watershed(int nhalo, const long long *off_h, const long long *coff_h,
                             const float *den, const long long *nbr, int k, const int *type,
                             const double *mass, const int *peak, const CoreOut &o,
                             const CatShellOut &so, long long small_max, long long *nshell_h,
                             int *herr_h, Workspace &ws, hipStream_t s);
// *err: 1 more shells than maxshell, 2 a root's own core never enclosed (two
// cores sharing a peak), 3 more core entries than maxcores.
hipError_t core_shells(const float *den, long long np, const long long *nbr, int k,
                       const int *peak, const float *coredensity, int numcore, const float *thr,
                       int nthr, unsigned char *flag, int *haloid, const ShellOut &o,
                       long long *nshell, int *err, Workspace &ws, hipStream_t s);

// ---- TSC deposit / read-back (tsc.hip) ----------------------------------
template <class P>
hipError_t tsc_deposit(const P *x, const P *y, const P *z, const P *m,
                       const int *type, long long np, int ptype, const Grid &g,
                       float *den, Workspace &ws, hipStream_t s);
// order (optional): particle i of x, y, z is particle order[i] of out --
// the read-back then runs over coordinates already in a spatial order.
template <class P>
hipError_t tsc_readback(const P *x, const P *y, const P *z, long long np,
                        const Grid &g, const float *den, float *out,
                        Workspace &ws, hipStream_t s, const unsigned *order = nullptr);
// Batched forms: every halo of `b` deposited into / read back from its own
// grid of the batch grid buffer `den` (HaloSeg::goff).  No host
// synchronisation, no memset: safe to capture in a HIP graph.
template <class P>
hipError_t tsc_deposit_batch(const P *x, const P *y, const P *z, const P *m,
                             const BatchDev &b, float *den, Workspace &ws,
                             hipStream_t s);
// order/shift (optional): the coordinates are in sorted slots -- particle
// i of halo h at slot i + shift[h] -- and slot p is particle order[p] of out.
template <class P>
hipError_t tsc_readback_batch(const P *x, const P *y, const P *z, const BatchDev &b,
                              const float *den, float *out, Workspace &ws,
                              hipStream_t s, const unsigned *order = nullptr,
                              const long long *shift = nullptr);
// Deposit tiles of one grid (16^3 nodes over the padded mx x ny x nz box).
long long tsc_tile_count(const Grid &g);
// Whether the deposit can take the grid (a tile holds at least one padded row).
bool tsc_grid_fits(const Grid &g);
// Deposit tiles one particle can touch in the grid, at most.
int tsc_tile_entries(const Grid &g);
// Adjoint of read-back and deposit (see tsc.hip): adds dL/dx, dL/dy, dL/dz,
// dL/dm of L(rho_tsc) to acc = [gx | gy | gz | gm] (FP64, np each) from
// gin = dL/drho_tsc, the forward smoothed grid F and the adjoint grid A.
template <class P>
hipError_t tsc_grad_readback(const P *x, const P *y, const P *z, const P *m, const float *gin,
                             long long np, const Grid &g, const float *F, const float *A,
                             double *acc, Workspace &ws, hipStream_t s);
template <class P>
hipError_t tsc_grad_readback_batch(const P *x, const P *y, const P *z, const P *m,
                                   const float *gin, const BatchDev &b, const float *F,
                                   const float *A, double *acc, Workspace &ws, hipStream_t s);
// Peak climbing (see tsc.hip): niter normalised gradient-ascent steps of
// `step` cells on the read-back field of the smoothed grid den, climbed
// positions to ox/oy/oz, and each particle's peak (the local maximum of den
// reached by integer steepest ascent from its climbed node; unpadded flat
// index within its halo's grid) to peak.
template <class P>
hipError_t peak_climb(const P *x, const P *y, const P *z, long long np, const Grid &g,
                      const float *den, int niter, double step, float *ox, float *oy,
                      float *oz, int *peak, hipStream_t s);
template <class P>
hipError_t peak_climb_batch(const P *x, const P *y, const P *z, const BatchDev &b,
                            const float *den, int niter, double step, float *ox, float *oy,
                            float *oz, int *peak, hipStream_t s);

// ---- Gaussian smoothing (smooth.hip) ------------------------------------
hipError_t gaussian_smooth(float *den, const Grid &g, float RG, Workspace &ws,
                           hipStream_t s);
// -(pi RG/h)^2 with RG/h rounded to float as gsmooth.c:67-69 does.
double smooth_expfac(float RG, double cellsize);
// Grids of one shape and one smoothing factor, smoothed together.
struct SmoothGroup {
    int nx, ny, nz;
    double expfac;
    const long long *goffs;   // device [count]: first float of each grid
    const long long *hgoffs;  // host copy (read while a plan is made)
    int count;
};
// Builds the group's FFT plans (host work and uploads; call before capture).
hipError_t smooth_group_prepare(const SmoothGroup &grp);
// The launch plan of a catalogue's smoothing (smooth.hip): axis passes of
// grids of different shapes with the same line length share one launch.
struct SmoothBatchPlan;
SmoothBatchPlan *smooth_batch_plan_create(const SmoothGroup *groups, int ngroups,
                                          hipError_t &err);
void smooth_batch_plan_free(SmoothBatchPlan *plan);
hipError_t gaussian_smooth_batch(float *den, const SmoothGroup *groups, int ngroups,
                                 const SmoothBatchPlan *plan, hipStream_t s);

// ---- kNN (knn.hip) -------------------------------------------------------
// Exact k nearest candidates of every query (self included when the query
// set is the candidate set), sorted by (FP64 dist2, candidate index).
// nbr (int64 [nq*k], candidate indices) and rho (float [nq]) are optional.
// rho_mode: 0 none, 1 unweighted W4 sum, 2 W4 weighted by candidate mass.
// cand_sel: optional int type array + ptype selecting the candidates among
// the same particle arrays (nullptr: every particle is a candidate).
// sorted_hook (optional) runs on the stream once the queries are in
// Hilbert order: FP64 copies of their coordinates and each sorted query's
// particle index (used by the density stage for a coherent read-back).
using SortedHook = std::function<hipError_t(const double *, const double *, const double *,
                                            const unsigned *, long long)>;
template <class P>
hipError_t knn(const P *x, const P *y, const P *z, const P *m, long long np,
               const int *type, int cand_ptype, int k, long long *nbr,
               float *rho, int rho_mode, Workspace &ws, hipStream_t s,
               const SortedHook *sorted_hook = nullptr);

// Device tables of a batched kNN, built once per catalogue (knn_batch_plan):
// per-halo trees in one box array (leaves of every halo first, in padded
// candidate order, then the internal levels level by level) and the wave
// table, largest halos first.
struct KnnBatchPlan {
    Schedule sched;  // fixed when the batch is created
    int nhalo = 0, k = 0, hbits = 0;
    int max_kb = 0;  // slot bits of the packed keys of the largest halo
    long long np = 0, npad = 0, nb_total = 0, nboxes = 0, nwaves = 0;
    void *halos = nullptr;  // KHalo[nhalo] (knn.hip)
    void *waves = nullptr;  // KWave[nwaves] (knn.hip)
    long long *child_begin = nullptr;      // [nboxes - nb_total] first child box
    unsigned char *child_cnt = nullptr;    // [nboxes - nb_total] children (1..8)
    long long *slot_shift = nullptr;       // [nhalo] sorted slot of a halo's particle i: i + shift
    std::vector<long long> level_begin;    // internal level l: boxes [level_begin[l-1], level_begin[l])
};
// off: host [nhalo+1] particle offsets (off[0] = 0, every halo >= k particles).
hipError_t knn_batch_plan(const long long *off, int nhalo, int k, KnnBatchPlan &plan);
void knn_batch_plan_free(KnnBatchPlan &plan);
// Exact k-NN lists (indices local to each halo) and W4 densities of every
// particle among the particles of its own halo; no host synchronisation.
// sorted_hook (optional): as for knn(), with the padded sorted slots
// (KnnBatchPlan::slot_shift maps particles to them).
template <class P>
hipError_t knn_batch(const P *x, const P *y, const P *z, const P *m, const BatchDev &b,
                     const KnnBatchPlan &plan, long long *nbr, float *rho, int rho_mode,
                     Workspace &ws, hipStream_t s, const SortedHook *sorted_hook = nullptr);

// Adjoint of the W4 kNN density over the forward lists nbr (np x k, indices
// into the same arrays): adds dL/dx.. and, when weighted, dL/dm of
// L(rho_knn) to acc = [gx | gy | gz | gm] (FP64 atomics), gin = dL/drho_knn.
template <class P>
hipError_t knn_density_backward(const P *x, const P *y, const P *z, const P *m, long long np,
                                const long long *nbr, int k, int weighted, const float *gin,
                                double *acc, Workspace &ws, hipStream_t s);
// Batched form: lists of the batch (indices local to each halo).
template <class P>
hipError_t knn_batch_backward(const P *x, const P *y, const P *z, const P *m, const BatchDev &b,
                              const KnnBatchPlan &plan, const long long *nbr, int weighted,
                              const float *gin, double *acc, Workspace &ws, hipStream_t s);
// out[i] = (float)acc[i] for the four gradient arrays (any may be NULL).
hipError_t grad_finalize(const double *acc, long long np, float *gx, float *gy, float *gz,
                         float *gm, hipStream_t s);

// AoS SimpleBasicParticleType (88 B) -> SoA double arrays on device.
hipError_t unpack_particles(const void *d_bp, long long np, double *x, double *y,
                            double *z, double *m, int *type, hipStream_t s);

// ---- integer peak stage (peaks.hip) --------------------------------------
// Grid cores of lagFindStellarCore (nnost.c:1115-1271): core particle
// indices, in (k, j, i) node order, to core[0 .. *ncore) (at most maxcore
// written; *ncore is a device word and may exceed maxcore).
template <class P>
hipError_t grid_cores(const P *x, const P *y, const P *z, const int *type, long long np,
                      int ptype, const float *densph, const Grid &g, const float *den,
                      double threshold, int nbuff, int *core, long long maxcore,
                      unsigned *ncore, Workspace &ws, hipStream_t s);
// finddenpeak (subhaloden.mod6.c:416-492): core particle indices in index
// order, as grid_cores.
hipError_t find_den_peak(const float *den, int k, const long long *nbr, long long np,
                         const int *type, int jflag, double threshold, int *core,
                         long long maxcore, unsigned *ncore, Workspace &ws, hipStream_t s);
// MergingPeak (subhaloden.mod6.c:341-415) over n cores: core i sits at the
// particle peak[i] of np (peak == nullptr: at x[i], y[i], z[i]) and is ranked by
// key[i], largest first.  keep[0 .. *nkeep): the input indices of the kept
// cores in rank order (*nkeep a device word); order_out (nullable): the input
// index of each rank.
template <class P>
hipError_t merging_peak(const P *x, const P *y, const P *z, long long np, const int *peak,
                        const float *key, int n, double link, int *keep, unsigned *nkeep,
                        int *order_out, Workspace &ws, hipStream_t s);

// ---- lagFindStellarCore's grid (capi.hip) ---------------------------------
constexpr double kLagCellSize = 0.004;  // TSC_CELL_SIZE (params.h:92)
// The TSC grid nnost.c:1040-1058 sizes for n particles (coordinates
// x[i * stride], ...): node counts and origin; the cell size is kLagCellSize.
void lag_grid_box(const double *x, const double *y, const double *z, long long n,
                  long long stride, int nxyz[3], double xyzmin[3]);

// ---- phase timing (capi.hip) ---------------------------------------------
void timing_mark(const char *phase, hipStream_t s, bool end);
// Phase timing is suspended on this thread while a graph is being captured.
void timing_suspend(bool on);

}  // namespace pgalf

// ---- status of the C-ABI entry points (capi.hip) -----------------------------
#include <mutex>
extern std::mutex g_api_mu;  // serialises entry points sharing a device workspace
int set_err(int code, const char *fmt, ...);
int ok();
int hip_err(hipError_t e, const char *where);
// Read and clear the workspace's device status word (synchronises s).
int consume_status(pgalf::Workspace &ws, hipStream_t s, const char *where);

#define PGALF_HIP_TRY(expr)                       \
    do {                                          \
        hipError_t _e = (expr);                   \
        if (_e != hipSuccess) return _e;          \
    } while (0)
