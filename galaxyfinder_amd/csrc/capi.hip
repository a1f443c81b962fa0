// C-ABI of the density-stage library: the reference's entry points
// (assign_density_TSC, gaussian_Smoothing, findDen, findsphdensity,
// starfindsphdensity; tsc_omp2.c / gsmooth.c / nnost.c) as synchronous
// host-array calls, plus the device-resident API used from Python.
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "pgalf_internal.h"

namespace pgalf {

// ---- workspace -------------------------------------------------------------
Workspace::Workspace(int device, void *base, size_t bytes) : device_(device), fixed_(true) {
    status_ = static_cast<unsigned *>(base);
    Chunk c;
    c.ptr = static_cast<char *>(base) + 256;
    c.cap = bytes > 256 ? bytes - 256 : 0;
    c.used = 0;
    chunks_.push_back(c);
}

Workspace::~Workspace() {
    if (fixed_) return;  // caller-owned memory
    for (auto &c : chunks_) hipFree(c.ptr);
    if (status_ && own_status_) hipFree(status_);
    if (done_) hipEventDestroy(done_);
}

static bool capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

void Workspace::end(hipStream_t s) {
    // under stream capture the ordering is the graph's own
    if (fixed_ || !done_ || capturing(s)) return;
    if (hipEventRecord(done_, s) == hipSuccess) {
        last_ = s;
        has_last_ = true;
    }
}

hipError_t Workspace::add_chunk(size_t bytes) {
    if (fixed_) return hipErrorOutOfMemory;
    Chunk c;
    c.cap = bytes;
    c.used = 0;
    hipError_t e = hipMalloc(&c.ptr, bytes);
    if (e != hipSuccess) return e;
    chunks_.push_back(c);
    return hipSuccess;
}

hipError_t Workspace::begin(hipStream_t s) {
    if (fixed_) {  // no events, no growth: reset the bump position only
        if (peak_ > high_water_) high_water_ = peak_;
        peak_ = 0;
        chunks_[0].used = 0;
        cur_ = 0;
        return hipSuccess;
    }
    if (!done_) PGALF_HIP_TRY(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    // the previous call ran on another stream: wait for its kernels
    if (has_last_ && s != last_ && !capturing(s)) PGALF_HIP_TRY(hipStreamWaitEvent(s, done_, 0));
    if (!status_) {
        PGALF_HIP_TRY(hipMalloc(&status_, sizeof(unsigned)));
        PGALF_HIP_TRY(hipMemsetAsync(status_, 0, sizeof(unsigned), s));
    }
    if (peak_ > high_water_) high_water_ = peak_;
    peak_ = 0;
    // several chunks (the last call grew) or one too small: one chunk of the
    // high-water size, allocated after the old ones are freed
    if (chunks_.size() > 1 || (!chunks_.empty() && chunks_[0].cap < high_water_)) {
        PGALF_HIP_TRY(hipStreamSynchronize(s));
        for (auto &c : chunks_) hipFree(c.ptr);
        chunks_.clear();
        PGALF_HIP_TRY(add_chunk(high_water_ + (high_water_ >> 3) + (1 << 20)));
    }
    for (auto &c : chunks_) c.used = 0;
    cur_ = 0;
    return hipSuccess;
}

void *Workspace::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    // the bump chunk, else a later one (emptied by a release), else a new one
    while (!chunks_.empty() && chunks_[cur_].used + bytes > chunks_[cur_].cap &&
           cur_ + 1 < chunks_.size())
        cur_++;
    if (chunks_.empty() || chunks_[cur_].used + bytes > chunks_[cur_].cap) {
        size_t cap = bytes < ((size_t)64 << 20) ? ((size_t)64 << 20) : bytes;
        if (add_chunk(cap) != hipSuccess) return nullptr;
        cur_ = chunks_.size() - 1;
    }
    Chunk &c = chunks_[cur_];
    void *p = c.ptr + c.used;
    c.used += bytes;
    size_t live = 0;
    for (size_t i = 0; i <= cur_; i++) live += chunks_[i].used;
    if (live > peak_) peak_ = live;
    return p;
}

void Workspace::release(const Mark &m) {
    if (chunks_.empty()) return;
    for (size_t i = m.chunk + 1; i < chunks_.size(); i++) chunks_[i].used = 0;
    chunks_[m.chunk].used = m.used;
    cur_ = m.chunk;
}

static std::mutex g_ws_mu;
static std::map<int, Workspace *> g_ws;

Workspace &workspace_for_current_device() {
    int dev = 0;
    hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_ws.find(dev);
    if (it != g_ws.end()) return *it->second;
    Workspace *w = new Workspace(dev);
    g_ws[dev] = w;
    return *w;
}

// ---- schedule ------------------------------------------------------------------
static std::mutex g_sched_mu;
static std::vector<Schedule> g_sched;  // bins by ascending nmax
static int g_smooth_direct_max = 56;

Schedule schedule_for(long long n) {
    std::lock_guard<std::mutex> lk(g_sched_mu);
    for (const Schedule &b : g_sched)
        if (n <= b.nmax) return b;
    return Schedule();
}

int smooth_direct_max() {
    std::lock_guard<std::mutex> lk(g_sched_mu);
    return g_smooth_direct_max;
}

// ---- timing ------------------------------------------------------------------
static std::mutex g_t_mu;
static bool g_t_on = false;
struct PendingEv {
    std::string name;
    hipEvent_t a, b;
    bool closed;
};
static std::vector<PendingEv> g_t_pending;
static std::map<std::string, std::pair<double, int>> g_t_total;

static thread_local bool t_t_suspended = false;
void timing_suspend(bool on) { t_t_suspended = on; }

void timing_mark(const char *phase, hipStream_t s, bool end) {
    std::lock_guard<std::mutex> lk(g_t_mu);
    if (!g_t_on || t_t_suspended) return;
    if (!end) {
        PendingEv p;
        p.name = phase;
        p.closed = false;
        hipEventCreate(&p.a);
        hipEventCreate(&p.b);
        hipEventRecord(p.a, s);
        g_t_pending.push_back(p);
    } else {
        for (auto it = g_t_pending.rbegin(); it != g_t_pending.rend(); ++it) {
            if (!it->closed && it->name == phase) {
                hipEventRecord(it->b, s);
                it->closed = true;
                break;
            }
        }
    }
}

static void timing_collect() {
    for (auto &p : g_t_pending) {
        if (p.closed) {
            hipEventSynchronize(p.b);
            float ms = 0.f;
            hipEventElapsedTime(&ms, p.a, p.b);
            auto &t = g_t_total[p.name];
            t.first += ms;
            t.second += 1;
        }
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    g_t_pending.clear();
}

// ---- AoS unpack ----------------------------------------------------------------
__global__ void unpack_k(const unsigned char *bp, long long np, double *x, double *y,
                         double *z, double *m, int *type) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const unsigned char *r = bp + i * sizeof(pgalf_particle_t);
    type[i] = (int)*reinterpret_cast<const unsigned *>(r + offsetof(pgalf_particle_t, type));
    m[i] = *reinterpret_cast<const double *>(r + offsetof(pgalf_particle_t, mass));
    x[i] = *reinterpret_cast<const double *>(r + offsetof(pgalf_particle_t, x));
    y[i] = *reinterpret_cast<const double *>(r + offsetof(pgalf_particle_t, y));
    z[i] = *reinterpret_cast<const double *>(r + offsetof(pgalf_particle_t, z));
}

hipError_t unpack_particles(const void *d_bp, long long np, double *x, double *y, double *z,
                            double *m, int *type, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    unpack_k<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(
        static_cast<const unsigned char *>(d_bp), np, x, y, z, m, type);
    return hipGetLastError();
}

// fn(lo, hi) over [0, n) split across up to 8 host threads (the calling
// one included; a thread that cannot start leaves its part to the caller)
template <class F>
static void par_for(size_t n, size_t min_serial, F fn) {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nt = hw < 2 ? 1u : (hw < 8 ? hw : 8u);
    if (nt == 1 || n < min_serial) {
        fn((size_t)0, n);
        return;
    }
    const size_t per = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned i = 1; i < nt && i * per < n; i++) {
        const size_t lo = i * per, hi = n - lo < per ? n : lo + per;
        try {
            th.emplace_back([&fn, lo, hi] { fn(lo, hi); });
        } catch (const std::system_error &) {
            fn(lo, hi);
        }
    }
    fn((size_t)0, per < n ? per : n);
    for (auto &t : th) t.join();
}

// Grid box of nnost.c:1040-1056: min/max of all particles, 1e-4 extent
// buffer on the max side, NCELLBUFF = 10 cells, origin shifted by 5 cells.
void lag_grid_box(const double *x, const double *y, const double *z, long long n,
                  long long stride, int nxyz[3], double xyzmin[3]) {
    const int nbuff = 10;  // NCELLBUFF (params.h:115)
    const double *c[3] = {x, y, z};
    // min and max over host threads (order-free: the same values)
    double plo[8][3], phi[8][3];
    std::atomic<int> part(0);
    par_for((size_t)(n > 0 ? n : 0), (size_t)1 << 16, [&](size_t a, size_t b) {
        double l[3] = {1e20, 1e20, 1e20}, h[3] = {-1e20, -1e20, -1e20};
        for (size_t i = a; i < b; i++)
            for (int d = 0; d < 3; d++) {
                l[d] = fmin(l[d], c[d][(long long)i * stride]);
                h[d] = fmax(h[d], c[d][(long long)i * stride]);
            }
        const int q = part.fetch_add(1);
        for (int d = 0; d < 3; d++) {
            plo[q][d] = l[d];
            phi[q][d] = h[d];
        }
    });
    for (int d = 0; d < 3; d++) {
        double lo = 1e20, hi = -1e20;
        for (int q = 0; q < part.load(); q++) {
            lo = fmin(lo, plo[q][d]);
            hi = fmax(hi, phi[q][d]);
        }
        hi = hi + (hi - lo) * 0.0001;
        nxyz[d] = (int)((hi - lo) / kLagCellSize + nbuff);
        xyzmin[d] = lo - kLagCellSize * nbuff / 2.;
    }
}

}  // namespace pgalf

using namespace pgalf;

static_assert(sizeof(pgalf_particle_t) == 88, "SimpleBasicParticleType is 88 bytes");
static_assert(offsetof(pgalf_particle_t, mass) == 8, "layout");
static_assert(offsetof(pgalf_particle_t, x) == 16, "layout");
static_assert(offsetof(pgalf_particle_t, indx) == 72, "layout");
static_assert(offsetof(pgalf_particle_t, bp) == 80, "layout");

// ---- status ----------------------------------------------------------------------
static thread_local int t_err = PGALF_OK;
static thread_local char t_msg[256] = "ok";
std::mutex g_api_mu;  // host entry points share one workspace per device

int set_err(int code, const char *fmt, ...) {
    t_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_msg, sizeof(t_msg), fmt, ap);
    va_end(ap);
    return code;
}
int ok() {
    t_err = PGALF_OK;
    snprintf(t_msg, sizeof(t_msg), "ok");
    return PGALF_OK;
}
int hip_err(hipError_t e, const char *where) {
    if (e == hipErrorOutOfMemory) return set_err(PGALF_ERR_NOMEM, "%s: out of device memory", where);
    if (e == hipErrorNotReady)
        return set_err(PGALF_ERR_SHORT_LIST, "%s: fewer candidates than neighbours", where);
    return set_err(PGALF_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

extern "C" int pgalf_last_error(void) { return t_err; }
extern "C" const char *pgalf_last_error_string(void) { return t_msg; }
extern "C" const char *pgalf_version(void) { return "pgalf-density-mi355x 0.1 (gfx950)"; }

extern "C" void pgalf_timing_enable(int on) {
    std::lock_guard<std::mutex> lk(g_t_mu);
    g_t_on = on != 0;
}
extern "C" void pgalf_timing_reset(void) {
    std::lock_guard<std::mutex> lk(g_t_mu);
    timing_collect();
    g_t_total.clear();
}
extern "C" int pgalf_timing_get(const char *phase, double *total_ms, int *count) {
    std::lock_guard<std::mutex> lk(g_t_mu);
    timing_collect();
    auto it = g_t_total.find(phase);
    if (it == g_t_total.end()) {
        if (total_ms) *total_ms = 0.0;
        if (count) *count = 0;
        return 0;
    }
    if (total_ms) *total_ms = it->second.first;
    if (count) *count = it->second.second;
    return 1;
}

// Read and clear the device status word (synchronises s).
int consume_status(Workspace &ws, hipStream_t s, const char *where) {
    unsigned st = 0;
    hipError_t e = hipMemcpyAsync(&st, ws.status(), sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemsetAsync(ws.status(), 0, sizeof(unsigned), s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(e, where);
    if (st & DEV_ERR_OUT_OF_GRID)
        return set_err(PGALF_ERR_OUT_OF_GRID, "%s: particles outside the grid were skipped", where);
    if (st & DEV_ERR_SHORT_LIST)
        return set_err(PGALF_ERR_SHORT_LIST, "%s: short neighbour list", where);
    if (st & DEV_ERR_BAD_INDEX)
        return set_err(PGALF_ERR_ARG, "%s: an index outside the particle range was skipped",
                       where);
    return ok();
}

// ---- host copies through pinned staging ------------------------------------
// The host-array entry points move the caller's pageable arrays (the particle
// records in, grids and neighbour lists out).  A copy from or to pageable
// memory runs well below the link's speed, so large copies go through two
// pinned staging buffers per device: the DMA of one chunk overlaps host
// threads copying the neighbouring chunk between the caller's memory and the
// other buffer.  Small copies go directly.  Both return with the caller's
// memory no longer needed (host to device) or filled (device to host).
constexpr size_t kStageBytes = (size_t)32 << 20;
constexpr size_t kStageMin = (size_t)8 << 20;
struct Staging {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
};

static hipError_t staging_for_current_device(Staging **out) {
    static std::map<int, Staging> st;  // under g_api_mu
    int dev = 0;
    PGALF_HIP_TRY(hipGetDevice(&dev));
    Staging &x = st[dev];
    for (int b = 0; b < 2; b++) {
        if (!x.buf[b]) PGALF_HIP_TRY(hipHostMalloc(&x.buf[b], kStageBytes, hipHostMallocDefault));
        if (!x.ev[b]) PGALF_HIP_TRY(hipEventCreateWithFlags(&x.ev[b], hipEventDisableTiming));
    }
    *out = &x;
    return hipSuccess;
}

static void par_memcpy(void *dst, const void *src, size_t n) {
    par_for(n, (size_t)1 << 20, [dst, src](size_t lo, size_t hi) {
        memcpy(static_cast<char *>(dst) + lo, static_cast<const char *>(src) + lo, hi - lo);
    });
}

static hipError_t host_to_dev(void *d, const void *h, size_t n, hipStream_t s) {
    if (n < kStageMin) return hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, s);
    Staging *st = nullptr;
    PGALF_HIP_TRY(staging_for_current_device(&st));
    for (size_t o = 0, i = 0; o < n; o += kStageBytes, i++) {
        const int b = (int)(i & 1);
        const size_t len = n - o < kStageBytes ? n - o : kStageBytes;
        // the buffer's previous DMA (of this call or an earlier one) is done
        PGALF_HIP_TRY(hipEventSynchronize(st->ev[b]));
        par_memcpy(st->buf[b], static_cast<const char *>(h) + o, len);
        PGALF_HIP_TRY(hipMemcpyAsync(static_cast<char *>(d) + o, st->buf[b], len,
                                     hipMemcpyHostToDevice, s));
        PGALF_HIP_TRY(hipEventRecord(st->ev[b], s));
    }
    return hipSuccess;
}

static hipError_t dev_to_host(void *h, const void *d, size_t n, hipStream_t s) {
    if (n < kStageMin) {
        PGALF_HIP_TRY(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, s));
        return hipStreamSynchronize(s);
    }
    Staging *st = nullptr;
    PGALF_HIP_TRY(staging_for_current_device(&st));
    const size_t nch = (n + kStageBytes - 1) / kStageBytes;
    for (size_t i = 0; i <= nch; i++) {
        if (i < nch) {  // chunk i into buffer i % 2 (after that buffer's earlier DMAs: stream order)
            const size_t o = i * kStageBytes, len = n - o < kStageBytes ? n - o : kStageBytes;
            PGALF_HIP_TRY(hipMemcpyAsync(st->buf[i & 1], static_cast<const char *>(d) + o, len,
                                         hipMemcpyDeviceToHost, s));
            PGALF_HIP_TRY(hipEventRecord(st->ev[i & 1], s));
        }
        if (i > 0) {  // chunk i - 1 out of its buffer while chunk i moves
            const size_t j = i - 1, o = j * kStageBytes;
            const size_t len = n - o < kStageBytes ? n - o : kStageBytes;
            PGALF_HIP_TRY(hipEventSynchronize(st->ev[j & 1]));
            par_memcpy(static_cast<char *>(h) + o, st->buf[j & 1], len);
        }
    }
    return hipSuccess;
}

static bool check_grid(int nx, int ny, int nz) {
    return nx >= 3 && ny >= 3 && nz >= 3 && (long long)nx * ny * nz < (1LL << 31);
}

// ---- host-array entry points (reference drop-in) ------------------------------------
namespace {

struct DeviceParticles {
    double *x, *y, *z, *m;
    int *type;
};

// The fields the stage reads (x, y, z, mass, type: 36 of the record's 88
// bytes) into device SoA arrays.  Large arrays are gathered on host threads
// into the pinned staging buffers, one chunk of SoA sections at a time, so
// the link carries only those fields; small ones go as records and are
// unpacked on the device.
hipError_t upload_particles(const pgalf_particle_t *bp, long long np, Workspace &ws,
                            hipStream_t s, DeviceParticles &d) {
    d.x = ws.alloc_n<double>(np);
    d.y = ws.alloc_n<double>(np);
    d.z = ws.alloc_n<double>(np);
    d.m = ws.alloc_n<double>(np);
    d.type = ws.alloc_n<int>(np);
    if (!d.x || !d.y || !d.z || !d.m || !d.type) return hipErrorOutOfMemory;
    if (sizeof(pgalf_particle_t) * (size_t)np < kStageMin) {
        void *raw = ws.alloc(sizeof(pgalf_particle_t) * (size_t)np);
        if (!raw) return hipErrorOutOfMemory;
        PGALF_HIP_TRY(hipMemcpyAsync(raw, bp, sizeof(pgalf_particle_t) * (size_t)np,
                                     hipMemcpyHostToDevice, s));
        return unpack_particles(raw, np, d.x, d.y, d.z, d.m, d.type, s);
    }
    Staging *st = nullptr;
    PGALF_HIP_TRY(staging_for_current_device(&st));
    const size_t per = kStageBytes / (4 * sizeof(double) + sizeof(int)) / 64 * 64;
    for (size_t o = 0, i = 0; o < (size_t)np; o += per, i++) {
        const int b = (int)(i & 1);
        const size_t c = (size_t)np - o < per ? (size_t)np - o : per;
        double *hx = static_cast<double *>(st->buf[b]);
        double *hy = hx + per, *hz = hy + per, *hm = hz + per;
        int *ht = reinterpret_cast<int *>(hm + per);
        PGALF_HIP_TRY(hipEventSynchronize(st->ev[b]));  // the buffer's previous DMAs are done
        const pgalf_particle_t *src = bp + o;
        par_for(c, (size_t)1 << 16, [=](size_t lo, size_t hi) {
            for (size_t j = lo; j < hi; j++) {
                hx[j] = src[j].x;
                hy[j] = src[j].y;
                hz[j] = src[j].z;
                hm[j] = src[j].mass;
                ht[j] = (int)src[j].type;
            }
        });
        PGALF_HIP_TRY(hipMemcpyAsync(d.x + o, hx, sizeof(double) * c, hipMemcpyHostToDevice, s));
        PGALF_HIP_TRY(hipMemcpyAsync(d.y + o, hy, sizeof(double) * c, hipMemcpyHostToDevice, s));
        PGALF_HIP_TRY(hipMemcpyAsync(d.z + o, hz, sizeof(double) * c, hipMemcpyHostToDevice, s));
        PGALF_HIP_TRY(hipMemcpyAsync(d.m + o, hm, sizeof(double) * c, hipMemcpyHostToDevice, s));
        PGALF_HIP_TRY(hipMemcpyAsync(d.type + o, ht, sizeof(int) * c, hipMemcpyHostToDevice, s));
        PGALF_HIP_TRY(hipEventRecord(st->ev[b], s));
    }
    return hipSuccess;
}

Grid make_grid(int nx, int ny, int nz, double x0, double y0, double z0, double h) {
    Grid g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.x0 = x0;
    g.y0 = y0;
    g.z0 = z0;
    g.h = h;
    return g;
}

}  // namespace

extern "C" void assign_density_TSC(pgalf_particle_t *bp, int np, float *den, int nx, int ny,
                                   int nz, double Xmin, double Ymin, double Zmin,
                                   double cellsize, int ptype) {
    if (np < 0 || !den || (np > 0 && !bp) || !check_grid(nx, ny, nz) || !(cellsize > 0)) {
        set_err(PGALF_ERR_ARG, "assign_density_TSC: bad arguments");
        return;
    }
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    Grid g = make_grid(nx, ny, nz, Xmin, Ymin, Zmin, cellsize);
    DeviceParticles d{};
    float *dden = nullptr;
    if (e == hipSuccess) e = upload_particles(bp, np, ws, s, d);
    if (e == hipSuccess) {
        dden = ws.alloc_n<float>(g.ncells());
        if (!dden) e = hipErrorOutOfMemory;
    }
    if (e == hipSuccess)
        e = tsc_deposit<double>(d.x, d.y, d.z, d.m, d.type, np, ptype, g, dden, ws, s);
    if (e == hipSuccess)
        e = dev_to_host(den, dden, sizeof(float) * g.ncells(), s);
    if (e != hipSuccess) {
        hip_err(e, "assign_density_TSC");
        return;
    }
    consume_status(ws, s, "assign_density_TSC");
}

extern "C" void gaussian_Smoothing(float *denGrid, int nx, int ny, int nz, double cellsize,
                                   float RG) {
    if (!denGrid || !check_grid(nx, ny, nz) || !(cellsize > 0)) {
        set_err(PGALF_ERR_ARG, "gaussian_Smoothing: bad arguments");
        return;
    }
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    Grid g = make_grid(nx, ny, nz, 0, 0, 0, cellsize);
    float *dden = nullptr;
    if (e == hipSuccess) {
        dden = ws.alloc_n<float>(g.ncells());
        if (!dden) e = hipErrorOutOfMemory;
    }
    if (e == hipSuccess)
        e = host_to_dev(dden, denGrid, sizeof(float) * g.ncells(), s);
    if (e == hipSuccess) e = gaussian_smooth(dden, g, RG, ws, s);
    if (e == hipSuccess) e = dev_to_host(denGrid, dden, sizeof(float) * g.ncells(), s);
    if (e != hipSuccess) {
        hip_err(e, "gaussian_Smoothing");
        return;
    }
    consume_status(ws, s, "gaussian_Smoothing");
}

extern "C" void findDen(pgalf_particle_t *bp, int np, float *densph, double xmin, double ymin,
                        double zmin, double cellsize, float *den, int nx, int ny, int nz) {
    if (np < 0 || !den || (np > 0 && (!bp || !densph)) || !check_grid(nx, ny, nz) ||
        !(cellsize > 0)) {
        set_err(PGALF_ERR_ARG, "findDen: bad arguments");
        return;
    }
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    Grid g = make_grid(nx, ny, nz, xmin, ymin, zmin, cellsize);
    DeviceParticles d{};
    float *dden = nullptr, *dout = nullptr;
    if (e == hipSuccess) e = upload_particles(bp, np, ws, s, d);
    if (e == hipSuccess) {
        dden = ws.alloc_n<float>(g.ncells());
        dout = ws.alloc_n<float>(np > 0 ? np : 1);
        if (!dden || !dout) e = hipErrorOutOfMemory;
    }
    if (e == hipSuccess)
        e = host_to_dev(dden, den, sizeof(float) * g.ncells(), s);
    if (e == hipSuccess) e = tsc_readback<double>(d.x, d.y, d.z, np, g, dden, dout, ws, s);
    if (e == hipSuccess && np > 0) e = dev_to_host(densph, dout, sizeof(float) * np, s);
    if (e != hipSuccess) {
        hip_err(e, "findDen");
        return;
    }
    consume_status(ws, s, "findDen");
}

// One kNN pass of host_knn: k neighbours among the candidates of type
// cand_ptype (PGALF_TYPE_ALL: every particle), lists and/or density.
struct HostKnnJob {
    int k;
    long long *nearindex;
    float *dens;
    int rho_mode;
    int cand_ptype;
};

// kNN lists (+ optional density) for host particles: the particles are
// uploaded once, then the jobs run in order, each one's outputs copied back
// before the next starts (a job that fails ends the call with its error).
static int host_knn(pgalf_particle_t *bp, int np, const HostKnnJob *jobs, int njobs,
                    const char *where) {
    auto check = [&](const HostKnnJob &jb) {
        if (np <= 0 || !bp || jb.k < 1 || jb.k > 32)
            return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
        if (jb.k > np)
            return set_err(PGALF_ERR_SHORT_LIST, "%s: Numnear %d > np %d", where, jb.k, np);
        return (int)PGALF_OK;
    };
    if (njobs < 1) return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    const int first = check(jobs[0]);
    if (first != PGALF_OK) return first;
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    DeviceParticles d{};
    if (e == hipSuccess) e = upload_particles(bp, np, ws, s, d);
    if (e != hipSuccess) return hip_err(e, where);
    for (int j = 0; j < njobs; j++) {
        const HostKnnJob &jb = jobs[j];
        const int ok_job = check(jb);
        if (ok_job != PGALF_OK) return ok_job;
        long long *dnbr = nullptr;
        float *drho = nullptr;
        if (jb.nearindex) {
            dnbr = ws.alloc_n<long long>((size_t)np * jb.k);
            if (!dnbr) e = hipErrorOutOfMemory;
        }
        if (e == hipSuccess && jb.dens) {
            drho = ws.alloc_n<float>(np);
            if (!drho) e = hipErrorOutOfMemory;
        }
        if (e == hipSuccess)
            e = knn<double>(d.x, d.y, d.z, d.m, np,
                            jb.cand_ptype == PGALF_TYPE_ALL ? nullptr : d.type, jb.cand_ptype,
                            jb.k, dnbr, drho, jb.rho_mode, ws, s);
        if (e == hipSuccess && dnbr)
            e = dev_to_host(jb.nearindex, dnbr, sizeof(long long) * (size_t)np * jb.k, s);
        if (e == hipSuccess && drho) e = dev_to_host(jb.dens, drho, sizeof(float) * np, s);
        if (e != hipSuccess) return hip_err(e, where);
        const int st = consume_status(ws, s, where);
        if (st != PGALF_OK) return st;
    }
    return PGALF_OK;
}

static int host_knn(pgalf_particle_t *bp, int np, int k, long long *nearindex, float *dens,
                    int rho_mode, int cand_ptype, const char *where) {
    const HostKnnJob job{k, nearindex, dens, rho_mode, cand_ptype};
    return host_knn(bp, np, &job, 1, where);
}

extern "C" void findsphdensity(pgalf_particle_t *bp, int np, long long *nearindex, int Numnear,
                               float *densph) {
    std::lock_guard<std::mutex> lk(g_api_mu);
    host_knn(bp, np, Numnear, nearindex, densph, 1, PGALF_TYPE_ALL, "findsphdensity");
}

extern "C" void starfindsphdensity(pgalf_particle_t *bp, int np, long long *nearindex,
                                   int Numnear, float *densph) {
    std::lock_guard<std::mutex> lk(g_api_mu);
    if (np <= 0 || !bp) {
        set_err(PGALF_ERR_ARG, "starfindsphdensity: bad arguments");
        return;
    }
    int nstar = 0;
    for (int i = 0; i < np; i++) nstar += bp[i].type == PGALF_TYPE_STAR;
    if (nstar == 0) {
        set_err(PGALF_ERR_SHORT_LIST, "starfindsphdensity: no star particles");
        return;
    }
    // NUMNEARDEN = 10 nearest stars (params.h:73); fewer when fewer stars exist,
    // as Find_Near returns them (nnost.c:698-703).
    int kden = nstar < 10 ? nstar : 10;
    // one upload for both passes: the star densities, then the lists
    HostKnnJob jobs[2];
    int nj = 0;
    if (densph) jobs[nj++] = {kden, nullptr, densph, 2, PGALF_TYPE_STAR};
    if (nearindex) jobs[nj++] = {Numnear, nearindex, nullptr, 0, PGALF_TYPE_ALL};
    if (nj) host_knn(bp, np, jobs, nj, "starfindsphdensity");
}

extern "C" int pgalf_near_lists(pgalf_particle_t *bp, int np, int Numnear, long long *nearindex) {
    std::lock_guard<std::mutex> lk(g_api_mu);
    if (!nearindex) return set_err(PGALF_ERR_ARG, "pgalf_near_lists: null output");
    return host_knn(bp, np, Numnear, nearindex, nullptr, 0, PGALF_TYPE_ALL, "pgalf_near_lists");
}

// The density stage's read-back, run by the kNN once its queries are in
// Hilbert order: over the sorted FP64 coordinates (the same values as the
// caller's, so the same bits), writing each result to its particle.
// Neighbouring threads then gather neighbouring grid nodes.
static SortedHook sorted_readback(const Grid &g, const float *den, float *out, Workspace &ws,
                                  hipStream_t s) {
    Workspace *w = &ws;
    return [g, den, out, w, s](const double *sx, const double *sy, const double *sz,
                               const unsigned *order, long long n) {
        timing_mark("findden", s, false);
        const hipError_t e = tsc_readback<double>(sx, sy, sz, n, g, den, out, *w, s, order);
        timing_mark("findden", s, true);
        return e;
    };
}

extern "C" int pgalf_lag_density_stage(pgalf_particle_t *bp, int np, int Numnear, int ptype,
                                       float *densph, long long *nearindex, float **grid_out,
                                       int *nxyz_out, double *xyzmin_out) {
    std::lock_guard<std::mutex> lk(g_api_mu);
    const char *where = "pgalf_lag_density_stage";
    if (np <= 0 || !bp || !densph || Numnear < 1 || Numnear > 32)
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    if (nearindex && Numnear > np)
        return set_err(PGALF_ERR_SHORT_LIST, "%s: Numnear %d > np %d", where, Numnear, np);
    const double cellsize = kLagCellSize;
    const float RG = 0.008f;  // Gaussian_Smoothing_Length (params.h:104)
    constexpr long long kStride = sizeof(pgalf_particle_t) / sizeof(double);
    int nn[3];
    double lo[3];
    lag_grid_box(&bp[0].x, &bp[0].y, &bp[0].z, np, kStride, nn, lo);
    const int nx = nn[0], ny = nn[1], nz = nn[2];
    const double xmin = lo[0], ymin = lo[1], zmin = lo[2];
    if (!check_grid(nx, ny, nz)) return set_err(PGALF_ERR_ARG, "%s: grid too large", where);
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    Grid g = make_grid(nx, ny, nz, xmin, ymin, zmin, cellsize);
    DeviceParticles d{};
    float *dden = nullptr, *drho = nullptr;
    long long *dnbr = nullptr;
    if (e == hipSuccess) e = upload_particles(bp, np, ws, s, d);
    if (e == hipSuccess) {
        dden = ws.alloc_n<float>(g.ncells());
        drho = ws.alloc_n<float>(np);
        if (nearindex) dnbr = ws.alloc_n<long long>((size_t)np * Numnear);
        if (!dden || !drho || (nearindex && !dnbr)) e = hipErrorOutOfMemory;
    }
    if (e == hipSuccess) e = tsc_deposit<double>(d.x, d.y, d.z, d.m, d.type, np, ptype, g, dden, ws, s);
    if (e == hipSuccess) e = gaussian_smooth(dden, g, RG, ws, s);
    if (e == hipSuccess && nearindex) {
        const SortedHook rb = sorted_readback(g, dden, drho, ws, s);
        e = knn<double>(d.x, d.y, d.z, d.m, np, nullptr, PGALF_TYPE_ALL, Numnear, dnbr, nullptr, 0,
                        ws, s, &rb);
    } else if (e == hipSuccess) {
        e = tsc_readback<double>(d.x, d.y, d.z, np, g, dden, drho, ws, s);
    }
    if (e == hipSuccess) e = dev_to_host(densph, drho, sizeof(float) * np, s);
    if (e == hipSuccess && nearindex)
        e = dev_to_host(nearindex, dnbr, sizeof(long long) * (size_t)np * Numnear, s);
    float *hgrid = nullptr;
    if (e == hipSuccess && grid_out) {
        hgrid = (float *)malloc(sizeof(float) * g.ncells());
        if (!hgrid) return set_err(PGALF_ERR_NOMEM, "%s: host allocation failed", where);
        e = dev_to_host(hgrid, dden, sizeof(float) * g.ncells(), s);
    }
    if (e != hipSuccess) {
        free(hgrid);
        return hip_err(e, where);
    }
    int rc = consume_status(ws, s, where);
    if (grid_out) *grid_out = hgrid;
    if (nxyz_out) { nxyz_out[0] = nx; nxyz_out[1] = ny; nxyz_out[2] = nz; }
    if (xyzmin_out) { xyzmin_out[0] = xmin; xyzmin_out[1] = ymin; xyzmin_out[2] = zmin; }
    return rc;
}

// ---- device-resident API ---------------------------------------------------------------
static Grid grid_from(const pgalf_grid_t *g) {
    return make_grid(g->nx, g->ny, g->nz, g->xmin, g->ymin, g->zmin, g->cellsize);
}

extern "C" int pgalf_dev_tsc(const float *x, const float *y, const float *z, const float *m,
                             const int *type, long long np, int ptype, const pgalf_grid_t *g,
                             float *den, void *stream) {
    if (!g || !den || !aligned8(den) || np < 0 || !check_grid(g->nx, g->ny, g->nz))
        return set_err(PGALF_ERR_ARG, "pgalf_dev_tsc: bad arguments");
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    timing_mark("tsc", s, false);
    if (e == hipSuccess) e = tsc_deposit<float>(x, y, z, m, type, np, ptype, grid_from(g), den, ws, s);
    timing_mark("tsc", s, true);
    return e == hipSuccess ? ok() : hip_err(e, "pgalf_dev_tsc");
}

extern "C" int pgalf_dev_smooth(float *den, const pgalf_grid_t *g, float RG, void *stream) {
    if (!g || !den || !aligned8(den) || !check_grid(g->nx, g->ny, g->nz))
        return set_err(PGALF_ERR_ARG, "pgalf_dev_smooth: bad arguments");
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    timing_mark("smooth", s, false);
    if (e == hipSuccess) e = gaussian_smooth(den, grid_from(g), RG, ws, s);
    timing_mark("smooth", s, true);
    return e == hipSuccess ? ok() : hip_err(e, "pgalf_dev_smooth");
}

extern "C" int pgalf_dev_findden(const float *x, const float *y, const float *z, long long np,
                                 const pgalf_grid_t *g, const float *den, float *densph,
                                 void *stream) {
    if (!g || !den || np < 0 || !check_grid(g->nx, g->ny, g->nz))
        return set_err(PGALF_ERR_ARG, "pgalf_dev_findden: bad arguments");
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    timing_mark("findden", s, false);
    if (e == hipSuccess) e = tsc_readback<float>(x, y, z, np, grid_from(g), den, densph, ws, s);
    timing_mark("findden", s, true);
    return e == hipSuccess ? ok() : hip_err(e, "pgalf_dev_findden");
}

extern "C" int pgalf_dev_knn(const float *x, const float *y, const float *z, const float *m,
                             const int *type, long long np, int cand_ptype, int k,
                             long long *nbr, float *rho, int rho_mode, void *stream) {
    if (np < 1 || k < 1 || k > 32 || rho_mode < 0 || rho_mode > 2 || (rho_mode == 2 && !m))
        return set_err(PGALF_ERR_ARG, "pgalf_dev_knn: bad arguments");
    if (k > np) return set_err(PGALF_ERR_SHORT_LIST, "pgalf_dev_knn: k > np");
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    timing_mark("knn", s, false);
    if (e == hipSuccess)
        e = knn<float>(x, y, z, m, np, type, type ? cand_ptype : PGALF_TYPE_ALL, k, nbr,
                       rho_mode ? rho : nullptr, rho_mode, ws, s);
    timing_mark("knn", s, true);
    return e == hipSuccess ? ok() : hip_err(e, "pgalf_dev_knn");
}

// Side stream of the one-halo density stage, per device (default
// priority), with a workspace of its own that reports into the device
// workspace's status word, and the events that order the two streams.
struct StageSide {
    hipStream_t s = nullptr;
    Workspace *ws = nullptr;
    hipEvent_t start = nullptr, sorted = nullptr, done = nullptr;
};

static hipError_t stage_side(Workspace &main_ws, StageSide **out) {
    static std::map<int, StageSide> sides;  // under g_api_mu
    StageSide &x = sides[main_ws.device()];
    if (!x.s) PGALF_HIP_TRY(hipStreamCreateWithFlags(&x.s, hipStreamNonBlocking));
    for (hipEvent_t *ev : {&x.start, &x.sorted, &x.done})
        if (!*ev) PGALF_HIP_TRY(hipEventCreateWithFlags(ev, hipEventDisableTiming));
    if (!x.ws) {
        x.ws = new Workspace(main_ws.device());
        x.ws->borrow_status(main_ws.status());
    }
    *out = &x;
    return hipSuccess;
}

extern "C" int pgalf_dev_density_stage(const float *x, const float *y, const float *z,
                                       const float *m, long long np, const pgalf_grid_t *g,
                                       float RG, int k, float *den, float *rho_tsc,
                                       long long *nbr, float *rho_knn, void *stream) {
    if (!g || !den || !aligned8(den) || np < 1 || k < 1 || k > 32 ||
        !check_grid(g->nx, g->ny, g->nz))
        return set_err(PGALF_ERR_ARG, "pgalf_dev_density_stage: bad arguments");
    if (k > np) return set_err(PGALF_ERR_SHORT_LIST, "pgalf_dev_density_stage: k > np");
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    Grid gg = grid_from(g);
    const bool with_knn = nbr || rho_knn;
    StageSide *side = nullptr;
    if (e == hipSuccess && with_knn && !capturing(s)) e = stage_side(ws, &side);
    if (e == hipSuccess && side) {
        // the grid chain on the side stream beside the kNN: the deposit and
        // smoothing while the kNN sorts and builds its tree, the read-back
        // (over the kNN's sorted coordinates) while it queries
        timing_mark("stage", s, false);
        e = hipEventRecord(side->start, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(side->s, side->start, 0);
        WsCall call2(*side->ws, side->s);
        if (e == hipSuccess) e = call2.e;
        timing_mark("tsc", side->s, false);
        if (e == hipSuccess)
            e = tsc_deposit<float>(x, y, z, m, nullptr, np, PGALF_TYPE_ALL, gg, den, *side->ws,
                                   side->s);
        timing_mark("tsc", side->s, true);
        timing_mark("smooth", side->s, false);
        if (e == hipSuccess) e = gaussian_smooth(den, gg, RG, *side->ws, side->s);
        timing_mark("smooth", side->s, true);
        SortedHook rb;
        if (rho_tsc) {
            const SortedHook on_side = sorted_readback(gg, den, rho_tsc, *side->ws, side->s);
            rb = [side, on_side, s](const double *sx, const double *sy, const double *sz,
                                    const unsigned *order, long long n) -> hipError_t {
                PGALF_HIP_TRY(hipEventRecord(side->sorted, s));
                PGALF_HIP_TRY(hipStreamWaitEvent(side->s, side->sorted, 0));
                return on_side(sx, sy, sz, order, n);
            };
        }
        timing_mark("knn", s, false);
        if (e == hipSuccess)
            e = knn<float>(x, y, z, m, np, nullptr, PGALF_TYPE_ALL, k, nbr, rho_knn,
                           rho_knn ? 1 : 0, ws, s, &rb);
        timing_mark("knn", s, true);
        // the caller's stream waits for the grid chain (also on an error:
        // the side stream's work must end before its scratch is reused)
        const hipError_t e2 = hipEventRecord(side->done, side->s);
        const hipError_t e3 = e2 == hipSuccess ? hipStreamWaitEvent(s, side->done, 0) : e2;
        if (e == hipSuccess) e = e3;
        if (e3 != hipSuccess) hipStreamSynchronize(side->s);
        timing_mark("stage", s, true);
        return e == hipSuccess ? ok() : hip_err(e, "pgalf_dev_density_stage");
    }
    timing_mark("stage", s, false);
    timing_mark("tsc", s, false);
    if (e == hipSuccess) e = tsc_deposit<float>(x, y, z, m, nullptr, np, PGALF_TYPE_ALL, gg, den, ws, s);
    timing_mark("tsc", s, true);
    timing_mark("smooth", s, false);
    if (e == hipSuccess) e = gaussian_smooth(den, gg, RG, ws, s);
    timing_mark("smooth", s, true);
    // with a kNN to run, the read-back waits for its Hilbert order (and is
    // timed inside the "knn" interval)
    if (e == hipSuccess && rho_tsc && !with_knn) {
        timing_mark("findden", s, false);
        e = tsc_readback<float>(x, y, z, np, gg, den, rho_tsc, ws, s);
        timing_mark("findden", s, true);
    }
    timing_mark("knn", s, false);
    if (e == hipSuccess && with_knn) {
        const SortedHook rb = rho_tsc ? sorted_readback(gg, den, rho_tsc, ws, s) : SortedHook();
        e = knn<float>(x, y, z, m, np, nullptr, PGALF_TYPE_ALL, k, nbr, rho_knn, rho_knn ? 1 : 0,
                       ws, s, &rb);
    }
    timing_mark("knn", s, true);
    timing_mark("stage", s, true);
    return e == hipSuccess ? ok() : hip_err(e, "pgalf_dev_density_stage");
}

extern "C" int pgalf_dev_density_stage_backward(
    const float *x, const float *y, const float *z, const float *m, long long np,
    const pgalf_grid_t *g, float RG, int k, const float *den, const long long *nbr,
    int rho_mode, const float *g_tsc, const float *g_knn, float *gx, float *gy, float *gz,
    float *gm, void *stream) {
    const char *where = "pgalf_dev_density_stage_backward";
    if (!g || np < 1 || !x || !y || !z || !m || !check_grid(g->nx, g->ny, g->nz) ||
        (g_tsc && !den) || (g_knn && (!nbr || k < 1 || k > 32 || k > np)) ||
        (g_knn && rho_mode != 1 && rho_mode != 2))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    const Grid gg = grid_from(g);
    double *acc = ws.alloc_n<double>(4 * (size_t)np);
    if (!acc) e = hipErrorOutOfMemory;
    timing_mark("bwd", s, false);
    if (e == hipSuccess) e = fill_u64(reinterpret_cast<unsigned long long *>(acc), 4 * np, 0ull, s);
    if (e == hipSuccess && g_tsc) {
        // adjoint grid A = S D_g (deposit of the read-back cotangent, smoothed:
        // the Gaussian filter is real and even, so S is self-adjoint)
        timing_mark("bwd_tsc", s, false);
        float *A = ws.alloc_n<float>(gg.ncells());
        if (!A) e = hipErrorOutOfMemory;
        if (e == hipSuccess)
            e = tsc_deposit<float>(x, y, z, g_tsc, nullptr, np, PGALF_TYPE_ALL, gg, A, ws, s);
        if (e == hipSuccess) e = gaussian_smooth(A, gg, RG, ws, s);
        if (e == hipSuccess) e = tsc_grad_readback<float>(x, y, z, m, g_tsc, np, gg, den, A, acc, ws, s);
        timing_mark("bwd_tsc", s, true);
    }
    if (e == hipSuccess && g_knn) {
        timing_mark("bwd_knn", s, false);
        e = knn_density_backward<float>(x, y, z, m, np, nbr, k, rho_mode == 2 ? 1 : 0, g_knn, acc,
                                        ws, s);
        timing_mark("bwd_knn", s, true);
    }
    if (e == hipSuccess) e = grad_finalize(acc, np, gx, gy, gz, gm, s);
    timing_mark("bwd", s, true);
    return e == hipSuccess ? ok() : hip_err(e, where);
}

extern "C" int pgalf_dev_peak_climb(const float *x, const float *y, const float *z,
                                    long long np, const pgalf_grid_t *g, const float *den,
                                    int niter, float step, float *ox, float *oy, float *oz,
                                    int *peak, void *stream) {
    const char *where = "pgalf_dev_peak_climb";
    if (!g || np < 0 || !den || !ox || !oy || !oz || !peak || niter < 0 || !(step >= 0.0f) ||
        (np > 0 && (!x || !y || !z)) || !check_grid(g->nx, g->ny, g->nz))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    if (e == hipSuccess)
        e = peak_climb<float>(x, y, z, np, grid_from(g), den, niter, (double)step, ox, oy, oz,
                              peak, s);
    return e == hipSuccess ? ok() : hip_err(e, where);
}

extern "C" int pgalf_dev_status(void *stream) {
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    if (!ws.status()) {
        WsCall call(ws, s);
        if (call.e != hipSuccess) return hip_err(call.e, "pgalf_dev_status");
    }
    return consume_status(ws, s, "pgalf_dev_status");
}

// ---- kNN with no allocation ------------------------------------------------------------
// Scratch of pgalf_dev_knn_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code never passes `scratch` itself
// to a kernel (in host code the name is a stand-in, not a device address):
// the first call on a device looks its address up with hipGetSymbolAddress
// and keeps it per device.
static constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_na_mu;
std::map<int, Workspace *> g_na_ws;  // per device: a fixed arena over its scratch copy

Workspace *noalloc_workspace(hipError_t &e) {
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_na_mu);
    auto it = g_na_ws.find(dev);
    if (it != g_na_ws.end()) return it->second;
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        if (e == hipSuccess) e = hipErrorInvalidSymbol;
        return nullptr;
    }
    Workspace *w = new Workspace(dev, ptr, kNoAllocScratchBytes);
    g_na_ws[dev] = w;
    return w;
}
}  // namespace

extern "C" long long pgalf_dev_knn_noalloc_max(void) { return PGALF_KNN_NOALLOC_MAX; }

extern "C" int pgalf_dev_knn_noalloc(const float *x, const float *y, const float *z,
                                     const float *m, const int *type, long long np,
                                     int cand_ptype, int k, long long *nbr, float *rho,
                                     int rho_mode, void *stream) {
    const char *where = "pgalf_dev_knn_noalloc";
    if (np < 1 || k < 1 || k > 32 || rho_mode < 0 || rho_mode > 2 || (rho_mode == 2 && !m))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    if (np > PGALF_KNN_NOALLOC_MAX)
        return set_err(PGALF_ERR_ARG, "%s: %lld particles exceed the %lld that its %zu-byte "
                       "static scratch holds (use pgalf_dev_knn)", where, np,
                       (long long)PGALF_KNN_NOALLOC_MAX, kNoAllocScratchBytes);
    if (k > np) return set_err(PGALF_ERR_SHORT_LIST, "%s: k > np", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipError_t e = hipSuccess;
    Workspace *ws = noalloc_workspace(e);
    if (!ws) return hip_err(e, where);
    hipStream_t s = (hipStream_t)stream;
    e = ws->begin(s);
    timing_suspend(true);  // phase timing would create events
    if (e == hipSuccess)
        e = knn<float>(x, y, z, m, np, type, type ? cand_ptype : PGALF_TYPE_ALL, k, nbr,
                       rho_mode ? rho : nullptr, rho_mode, *ws, s);
    timing_suspend(false);
    if (e == hipErrorOutOfMemory)
        return set_err(PGALF_ERR_NOMEM, "%s: the %zu-byte static scratch is too small",
                       where, kNoAllocScratchBytes);
    return e == hipSuccess ? ok() : hip_err(e, where);
}

// ---- integer peak stage ------------------------------------------------------------------
// Copies the cores' particle indices back and fills the reference-layout
// records (peak, its position as floats, its density; other fields zero).
static int fill_cores(const int *dcore, unsigned ncore_dev, int maxcore,
                      const pgalf_particle_t *bp, const float *dens, pgalf_core_t *core,
                      int *numcore, hipStream_t s, const char *where) {
    const long long nout = ncore_dev < (unsigned)maxcore ? (long long)ncore_dev : maxcore;
    std::vector<int> idx((size_t)(nout > 0 ? nout : 1));
    if (nout > 0) {
        hipError_t e = hipMemcpyAsync(idx.data(), dcore, sizeof(int) * nout,
                                      hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_err(e, where);
    }
    for (long long c = 0; c < nout; c++) {
        pgalf_core_t r;
        memset(&r, 0, sizeof(r));
        const int p = idx[c];
        r.peak = p;
        r.cx = (float)bp[p].x;
        r.cy = (float)bp[p].y;
        r.cz = (float)bp[p].z;
        r.density = dens[p];
        core[c] = r;
    }
    if (numcore) *numcore = (int)ncore_dev;
    if (ncore_dev > (unsigned)maxcore)
        return set_err(PGALF_ERR_CORE_OVERFLOW, "%s: %u cores, the array holds %d", where,
                       ncore_dev, maxcore);
    return ok();
}

extern "C" int pgalf_grid_cores(pgalf_particle_t *bp, int np, const float *densph,
                                const float *denGrid, int nx, int ny, int nz, double xmin,
                                double ymin, double zmin, double cellsize, int ptype,
                                float threshold, int nbuff, pgalf_core_t *core, int maxcore,
                                int *numcore) {
    const char *where = "pgalf_grid_cores";
    if (np < 0 || (np > 0 && (!bp || !densph)) || !denGrid || !core || maxcore < 0 ||
        nbuff < 2 || !check_grid(nx, ny, nz) || !(cellsize > 0))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    const Grid g = make_grid(nx, ny, nz, xmin, ymin, zmin, cellsize);
    DeviceParticles d{};
    if (e == hipSuccess && np > 0) e = upload_particles(bp, np, ws, s, d);
    float *dden = ws.alloc_n<float>(g.ncells()), *drho = ws.alloc_n<float>(np > 0 ? np : 1);
    int *dcore = ws.alloc_n<int>(maxcore > 0 ? maxcore : 1);
    unsigned *dn = ws.alloc_n<unsigned>(1);
    if (e == hipSuccess && (!dden || !drho || !dcore || !dn)) e = hipErrorOutOfMemory;
    if (e == hipSuccess)
        e = hipMemcpyAsync(dden, denGrid, sizeof(float) * g.ncells(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && np > 0)
        e = hipMemcpyAsync(drho, densph, sizeof(float) * np, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = grid_cores<double>(d.x, d.y, d.z, d.type, np, ptype, drho, g, dden,
                               (double)threshold, nbuff, dcore, maxcore, dn, ws, s);
    unsigned nc = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nc, dn, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(e, where);
    return fill_cores(dcore, nc, maxcore, bp, densph, core, numcore, s, where);
}

extern "C" int pgalf_finddenpeak(float *den, int numneigh, long long *neighbor, int np,
                                 int jflag, pgalf_particle_t *bp, pgalf_core_t *core,
                                 int maxcore, int *numcore) {
    const char *where = "pgalf_finddenpeak";
    if (np < 0 || (np > 0 && (!den || !neighbor || !bp)) || numneigh < 1 || !core ||
        maxcore < 0)
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    DeviceParticles d{};
    if (e == hipSuccess && np > 0) e = upload_particles(bp, np, ws, s, d);
    float *dden = ws.alloc_n<float>(np > 0 ? np : 1);
    long long *dnb = ws.alloc_n<long long>((size_t)(np > 0 ? np : 1) * numneigh);
    int *dcore = ws.alloc_n<int>(maxcore > 0 ? maxcore : 1);
    unsigned *dn = ws.alloc_n<unsigned>(1);
    if (e == hipSuccess && (!dden || !dnb || !dcore || !dn)) e = hipErrorOutOfMemory;
    if (e == hipSuccess && np > 0) {
        e = hipMemcpyAsync(dden, den, sizeof(float) * np, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dnb, neighbor, sizeof(long long) * (size_t)np * numneigh,
                               hipMemcpyHostToDevice, s);
    }
    if (e == hipSuccess)
        e = find_den_peak(dden, numneigh, dnb, np, d.type, jflag, (double)PGALF_PEAKTHRESHOLD,
                          dcore, maxcore, dn, ws, s);
    unsigned nc = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nc, dn, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(e, where);
    return fill_cores(dcore, nc, maxcore, bp, den, core, numcore, s, where);
}

extern "C" int pgalf_dev_grid_cores(const float *x, const float *y, const float *z,
                                    const int *type, long long np, int ptype,
                                    const float *densph, const pgalf_grid_t *g, const float *den,
                                    float threshold, int nbuff, int *core, long long maxcore,
                                    int *ncore, void *stream) {
    const char *where = "pgalf_dev_grid_cores";
    if (!g || !den || !core || !ncore || np < 0 || maxcore < 0 || nbuff < 2 ||
        (np > 0 && (!x || !y || !z || !densph)) || (ptype == PGALF_TYPE_STAR && !type) ||
        !check_grid(g->nx, g->ny, g->nz))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    if (e == hipSuccess)
        e = grid_cores<float>(x, y, z, type, np, ptype, densph, grid_from(g), den,
                              (double)threshold, nbuff, core, maxcore,
                              reinterpret_cast<unsigned *>(ncore), ws, s);
    return e == hipSuccess ? ok() : hip_err(e, where);
}

extern "C" int pgalf_dev_finddenpeak(const float *den, int k, const long long *nbr,
                                     long long np, const int *type, int jflag, float threshold,
                                     int *core, long long maxcore, int *ncore, void *stream) {
    const char *where = "pgalf_dev_finddenpeak";
    if (np < 0 || k < 1 || !core || !ncore || maxcore < 0 || (np > 0 && (!den || !nbr)) ||
        (jflag == 1 && !type))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    if (e == hipSuccess)
        e = find_den_peak(den, k, nbr, np, type, jflag, (double)threshold, core, maxcore,
                          reinterpret_cast<unsigned *>(ncore), ws, s);
    return e == hipSuccess ? ok() : hip_err(e, where);
}

extern "C" int pgalf_merging_peak(pgalf_particle_t *bp, int np, pgalf_core_t *core, int numcore,
                                  int iflag) {
    const char *where = "pgalf_merging_peak";
    if (numcore <= 0) {
        ok();
        return 0;
    }
    if (!bp || !core || np < 1 || numcore >= (1 << 30) || (iflag != 0 && iflag != 1)) {
        set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
        return -1;
    }
    // the cores' peak positions and keys, gathered on the host (FP64, as
    // the reference reads bp[core[i].peak])
    std::vector<double> pos(3 * (size_t)numcore);
    std::vector<float> key(numcore);
    for (int i = 0; i < numcore; i++) {
        const int p = core[i].peak;
        if (p < 0 || p >= np) {
            set_err(PGALF_ERR_ARG, "%s: core %d has peak %d outside [0, %d)", where, i, p, np);
            return -1;
        }
        pos[i] = bp[p].x;
        pos[numcore + i] = bp[p].y;
        pos[2 * (size_t)numcore + i] = bp[p].z;
        key[i] = iflag == 0 ? core[i].density : (float)core[i].numstar;
    }
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    double *dpos = ws.alloc_n<double>(3 * (size_t)numcore);
    float *dkey = ws.alloc_n<float>(numcore);
    int *dkeep = ws.alloc_n<int>(numcore), *dorder = ws.alloc_n<int>(numcore);
    unsigned *dn = ws.alloc_n<unsigned>(1);
    if (e == hipSuccess && (!dpos || !dkey || !dkeep || !dorder || !dn)) e = hipErrorOutOfMemory;
    if (e == hipSuccess)
        e = hipMemcpyAsync(dpos, pos.data(), sizeof(double) * pos.size(), hipMemcpyHostToDevice,
                           s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dkey, key.data(), sizeof(float) * numcore, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = merging_peak<double>(dpos, dpos + numcore, dpos + 2 * (size_t)numcore, numcore,
                                 nullptr, dkey, numcore, (double)PGALF_MERGINGPEAKLENGTH, dkeep,
                                 dn, dorder, ws, s);
    std::vector<int> keep(numcore), order(numcore);
    unsigned nk = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nk, dn, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(keep.data(), dkeep, sizeof(int) * numcore, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(order.data(), dorder, sizeof(int) * numcore, hipMemcpyDeviceToHost,
                           s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        hip_err(e, where);
        return -1;
    }
    // kept cores in rank order, then the ranked array past them
    std::vector<pgalf_core_t> src(core, core + numcore);
    for (unsigned j = 0; j < nk; j++) core[j] = src[keep[j]];
    for (int r = (int)nk; r < numcore; r++) core[r] = src[order[r]];
    ok();
    return (int)nk;
}

extern "C" int pgalf_dev_merging_peak(const float *x, const float *y, const float *z,
                                      long long np, const int *peak, const float *key,
                                      int numcore, float link, int *keep, int *order,
                                      int *nkeep, void *stream) {
    const char *where = "pgalf_dev_merging_peak";
    if (np < 1 || !x || !y || !z || !peak || !key || !keep || !nkeep || numcore < 0 ||
        numcore >= (1 << 30) || !(link >= 0.0f))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    if (e == hipSuccess)
        e = merging_peak<float>(x, y, z, np, peak, key, numcore, (double)link, keep,
                                reinterpret_cast<unsigned *>(nkeep), order, ws, s);
    return e == hipSuccess ? ok() : hip_err(e, where);
}

// ---- watershed of subhalo_den ------------------------------------------------------------
namespace {

int shell_err(int err, const char *where) {
    if (err == 2)
        return set_err(PGALF_ERR_ARG,
                       "%s: a root's own core was never enclosed (two cores share a peak)",
                       where);
    if (err == 1)
        return set_err(PGALF_ERR_CORE_OVERFLOW, "%s: more shells than maxshell", where);
    return set_err(PGALF_ERR_CORE_OVERFLOW, "%s: more shell core entries than maxcores", where);
}

template <class T>
hipError_t upload(T **d, const T *h, size_t n, Workspace &ws, hipStream_t s) {
    *d = ws.alloc_n<T>(n > 0 ? n : 1);
    if (!*d) return hipErrorOutOfMemory;
    return n > 0 ? hipMemcpyAsync(*d, h, sizeof(T) * n, hipMemcpyHostToDevice, s) : hipSuccess;
}

}  // namespace

extern "C" int pgalf_find_core_density(pgalf_particle_t *bp, int np, long long *neighbor,
                                       int NumNeighbor, pgalf_core_t *core, int numcore,
                                       const float *den, int *haloid, unsigned char *flag) {
    const char *where = "pgalf_find_core_density";
    if (np < 0 || np >= (1 << 26) || numcore < 0 || NumNeighbor < 1 || NumNeighbor > 32 ||
        (np > 0 && (!bp || !neighbor || !den || !haloid || !flag)) || (numcore > 0 && !core)) {
        set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
        return -1;
    }
    std::vector<int> peak(numcore);
    for (int c = 0; c < numcore; c++) {
        peak[c] = core[c].peak;
        if (peak[c] < 0 || peak[c] >= np) {
            set_err(PGALF_ERR_ARG, "%s: core %d has peak %d outside [0, %d)", where, c, peak[c],
                    np);
            return -1;
        }
    }
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    DeviceParticles d{};
    float *dden = nullptr;
    long long *dnb = nullptr;
    int *dpeak = nullptr;
    if (e == hipSuccess && np > 0) e = upload_particles(bp, np, ws, s, d);
    if (e == hipSuccess) e = upload(&dden, den, np, ws, s);
    if (e == hipSuccess) e = upload(&dnb, neighbor, (size_t)np * NumNeighbor, ws, s);
    if (e == hipSuccess) e = upload(&dpeak, peak.data(), numcore, ws, s);
    const int nc1 = numcore > 0 ? numcore : 1;
    CoreOut o;
    o.coredensity = ws.alloc_n<float>(nc1);
    o.nummem = ws.alloc_n<int>(nc1);
    o.numstar = ws.alloc_n<int>(nc1);
    o.starmass = ws.alloc_n<float>(nc1);
    o.haloid = ws.alloc_n<int>(np > 0 ? np : 1);
    o.flag = ws.alloc_n<unsigned char>(np > 0 ? np : 1);
    if (e == hipSuccess && (!o.coredensity || !o.nummem || !o.numstar || !o.starmass ||
                            !o.haloid || !o.flag))
        e = hipErrorOutOfMemory;
    int fail = 0;
    if (e == hipSuccess)
        e = find_core_density(dden, np, dnb, NumNeighbor, d.type, d.m, dpeak, numcore, o, &fail,
                              ws, s);
    std::vector<float> cd(nc1), sm(nc1);
    std::vector<int> nm(nc1), ns(nc1);
    if (e == hipSuccess && numcore > 0) {
        e = hipMemcpyAsync(cd.data(), o.coredensity, sizeof(float) * numcore,
                           hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(sm.data(), o.starmass, sizeof(float) * numcore,
                               hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(nm.data(), o.nummem, sizeof(int) * numcore, hipMemcpyDeviceToHost,
                               s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ns.data(), o.numstar, sizeof(int) * numcore,
                               hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess && np > 0) {
        e = hipMemcpyAsync(haloid, o.haloid, sizeof(int) * np, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(flag, o.flag, np, hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        hip_err(e, where);
        return -1;
    }
    if (consume_status(ws, s, where) != PGALF_OK) return -1;
    if (fail) {
        set_err(PGALF_ERR_ARG, "%s: a core-density search did not converge", where);
        return -1;
    }
    for (int c = 0; c < numcore; c++) {
        core[c].coredensity = cd[c];
        core[c].nummem = nm[c];
        core[c].numstar = ns[c];
        core[c].starmass = sm[c];
    }
    ok();
    return numcore;
}

extern "C" int pgalf_shell_thresholds(const float *den, int np, const pgalf_core_t *core,
                                      int numcore, float *thr) {
    const char *where = "pgalf_shell_thresholds";
    if (np < 0 || numcore < 0 || (np > 0 && !den) || (numcore > 0 && !core) || !thr) {
        set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
        return -1;
    }
    std::vector<float> cd(numcore);
    for (int c = 0; c < numcore; c++) cd[c] = core[c].coredensity;
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    float *dden = nullptr, *dcd = nullptr;
    if (e == hipSuccess) e = upload(&dden, den, np, ws, s);
    if (e == hipSuccess) e = upload(&dcd, cd.data(), numcore, ws, s);
    int n = 0;
    if (e == hipSuccess) e = shell_thresholds(dden, np, dcd, numcore, thr, &n, ws, s);
    if (e != hipSuccess) {
        hip_err(e, where);
        return -1;
    }
    ok();
    return n;
}

namespace {

int shells_run(const float *dden, long long np, const long long *dnb, int k, const int *dpeak,
               const float *dcd, int numcore, const float *thr, int nthr, unsigned char *dflag,
               int *dhalo, pgalf_shells_t *sh, bool host_shells, Workspace &ws, hipStream_t s,
               const char *where) {
    ShellOut o;
    o.maxshell = sh->maxshell;
    o.maxcores = sh->maxcores;
    hipError_t e = hipSuccess;
    if (host_shells) {
        o.part_ptr = ws.alloc_n<long long>(sh->maxshell + 1);
        o.core_ptr = ws.alloc_n<long long>(sh->maxshell + 1);
        o.part = ws.alloc_n<int>(np > 0 ? np : 1);
        o.cores = ws.alloc_n<int>(sh->maxcores > 0 ? sh->maxcores : 1);
        if (!o.part_ptr || !o.core_ptr || !o.part || !o.cores) e = hipErrorOutOfMemory;
    } else {
        o.part_ptr = sh->part_ptr;
        o.core_ptr = sh->core_ptr;
        o.part = sh->part;
        o.cores = sh->cores;
    }
    long long nshell = 0;
    int err = 0;
    if (e == hipSuccess)
        e = core_shells(dden, np, dnb, k, dpeak, dcd, numcore, thr, nthr, dflag, dhalo, o,
                        &nshell, &err, ws, s);
    if (e == hipSuccess && err == 0 && host_shells) {
        e = hipMemcpyAsync(sh->part_ptr, o.part_ptr, sizeof(long long) * (nshell + 1),
                           hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(sh->core_ptr, o.core_ptr, sizeof(long long) * (nshell + 1),
                               hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess && sh->part_ptr[nshell] > 0)
            e = hipMemcpyAsync(sh->part, o.part, sizeof(int) * sh->part_ptr[nshell],
                               hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && sh->core_ptr[nshell] > 0)
            e = hipMemcpyAsync(sh->cores, o.cores, sizeof(int) * sh->core_ptr[nshell],
                               hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(e, where);
    const int st = consume_status(ws, s, where);
    if (st != PGALF_OK) return st;
    if (err) return shell_err(err, where);
    sh->nshell = nshell;
    return ok();
}

}  // namespace

extern "C" int pgalf_shell_particles_peaks(int np, long long *neighbor, int NumNeighbor,
                                           pgalf_core_t *core, int numcore, const float *den,
                                           int *haloid, unsigned char *flag, const float *thr,
                                           int nthr, pgalf_shells_t *shells) {
    const char *where = "pgalf_shell_particles_peaks";
    if (np < 0 || np >= (1 << 26) || numcore < 0 || NumNeighbor < 1 || NumNeighbor > 32 ||
        nthr < 0 ||
        (nthr > 0 && !thr) || !shells || !shells->part_ptr || !shells->core_ptr ||
        !shells->part || !shells->cores || shells->maxshell < 0 || shells->maxcores < 0 ||
        (np > 0 && (!neighbor || !den || !haloid || !flag)) || (numcore > 0 && !core))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::vector<int> peak(numcore);
    std::vector<float> cd(numcore);
    for (int c = 0; c < numcore; c++) {
        peak[c] = core[c].peak;
        cd[c] = core[c].coredensity;
        if (peak[c] < 0 || peak[c] >= np)
            return set_err(PGALF_ERR_ARG, "%s: core %d has peak %d outside [0, %d)", where, c,
                           peak[c], np);
    }
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = nullptr;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    float *dden = nullptr, *dcd = nullptr;
    long long *dnb = nullptr;
    int *dpeak = nullptr, *dhalo = nullptr;
    unsigned char *dflag = nullptr;
    if (e == hipSuccess) e = upload(&dden, den, np, ws, s);
    if (e == hipSuccess) e = upload(&dnb, neighbor, (size_t)np * NumNeighbor, ws, s);
    if (e == hipSuccess) e = upload(&dpeak, peak.data(), numcore, ws, s);
    if (e == hipSuccess) e = upload(&dcd, cd.data(), numcore, ws, s);
    if (e == hipSuccess) e = upload(&dhalo, haloid, np, ws, s);
    if (e == hipSuccess) e = upload(&dflag, flag, np, ws, s);
    if (e != hipSuccess) return hip_err(e, where);
    const int r = shells_run(dden, np, dnb, NumNeighbor, dpeak, dcd, numcore, thr, nthr, dflag,
                             dhalo, shells, true, ws, s, where);
    if (r != PGALF_OK) return r;
    if (np > 0) {
        e = hipMemcpyAsync(haloid, dhalo, sizeof(int) * np, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(flag, dflag, np, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_err(e, where);
    }
    return ok();
}

extern "C" int pgalf_dev_find_core_density(const float *den, long long np, const long long *nbr,
                                           int k, const int *type, const double *mass,
                                           const int *peak, int numcore, float *coredensity,
                                           int *nummem, int *numstar, float *starmass,
                                           int *haloid, unsigned char *flag, void *stream) {
    const char *where = "pgalf_dev_find_core_density";
    if (np < 0 || np >= (1LL << 26) || numcore < 0 || k < 1 || k > 32 ||
        (np > 0 && (!den || !nbr || !type || !mass || !haloid || !flag)) ||
        (numcore > 0 && (!peak || !coredensity || !nummem || !numstar || !starmass)))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    CoreOut o{coredensity, nummem, numstar, starmass, haloid, flag};
    int fail = 0;
    if (e == hipSuccess)
        e = find_core_density(den, np, nbr, k, type, mass, peak, numcore, o, &fail, ws, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(e, where);
    const int st = consume_status(ws, s, where);
    if (st != PGALF_OK) return st;
    if (fail) return set_err(PGALF_ERR_ARG, "%s: a core-density search did not converge", where);
    return ok();
}

extern "C" int pgalf_dev_shell_thresholds(const float *den, long long np,
                                          const float *coredensity, int numcore, float *thr,
                                          int *nthr, void *stream) {
    const char *where = "pgalf_dev_shell_thresholds";
    if (np < 0 || numcore < 0 || (np > 0 && !den) || (numcore > 0 && !coredensity) || !thr ||
        !nthr)
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    if (e == hipSuccess) e = shell_thresholds(den, np, coredensity, numcore, thr, nthr, ws, s);
    return e == hipSuccess ? ok() : hip_err(e, where);
}

extern "C" int pgalf_dev_shell_particles_peaks(const float *den, long long np,
                                               const long long *nbr, int k, const int *peak,
                                               const float *coredensity, int numcore,
                                               const float *thr, int nthr, unsigned char *flag,
                                               int *haloid, pgalf_shells_t *shells,
                                               void *stream) {
    const char *where = "pgalf_dev_shell_particles_peaks";
    if (np < 0 || np >= (1LL << 26) || numcore < 0 || k < 1 || k > 32 || nthr < 0 ||
        (nthr > 0 && !thr) || !shells || !shells->part_ptr || !shells->core_ptr ||
        !shells->part || !shells->cores || shells->maxshell < 0 || shells->maxcores < 0 ||
        (np > 0 && (!den || !nbr || !flag || !haloid)) ||
        (numcore > 0 && (!peak || !coredensity)))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    if (call.e != hipSuccess) return hip_err(call.e, where);
    return shells_run(den, np, nbr, k, peak, coredensity, numcore, thr, nthr, flag, haloid,
                      shells, false, ws, s, where);
}

extern "C" long long pgalf_catalog_shell_layout(int nhalo, const long long *off,
                                                const long long *core_off, long long *shell_base,
                                                long long *core_base, long long *core_cap,
                                                long long *total_cores) {
    if (nhalo < 0 || !off || !core_off || (nhalo > 0 && (!shell_base || !core_base || !core_cap)) ||
        !total_cores) {
        set_err(PGALF_ERR_ARG, "pgalf_catalog_shell_layout: bad arguments");
        return -1;
    }
    long long tot_s = 0;
    catalog_shell_layout(nhalo, off, core_off, shell_base, core_base, core_cap, &tot_s,
                         total_cores);
    ok();
    return tot_s;
}

extern "C" int pgalf_dev_catalog_watershed(int nhalo, const long long *off,
                                           const long long *core_off, const float *den,
                                           const long long *nbr, int k, const int *type,
                                           const double *mass, const int *peak,
                                           float *coredensity, int *nummem, int *numstar,
                                           float *starmass, int *haloid, unsigned char *flag,
                                           long long *part_ptr, int *part, long long *core_ptr,
                                           int *cores, long long small_max, long long *nshell,
                                           int *halo_status, void *stream) {
    const char *where = "pgalf_dev_catalog_watershed";
    if (nhalo < 0 || !off || !core_off || k < 1 || k > 32 || small_max < 0 ||
        (nhalo > 0 && (!nshell || !halo_status)))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    if (off[0] != 0 || core_off[0] != 0)
        return set_err(PGALF_ERR_ARG, "%s: off[0] and core_off[0] must be 0", where);
    for (int h = 0; h < nhalo; h++)
        if (off[h + 1] < off[h] || core_off[h + 1] < core_off[h] ||
            off[h + 1] - off[h] >= (1LL << 26))
            return set_err(PGALF_ERR_ARG, "%s: halo %d has bad offsets or 2^26 particles", where,
                           h);
    const long long np = nhalo > 0 ? off[nhalo] : 0, nc = nhalo > 0 ? core_off[nhalo] : 0;
    if ((np > 0 && (!den || !nbr || !type || !mass || !haloid || !flag || !part)) ||
        (nc > 0 && (!peak || !coredensity || !nummem || !numstar || !starmass || !cores)) ||
        (nhalo > 0 && (!part_ptr || !core_ptr)))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    Workspace &ws = workspace_for_current_device();
    WsCall call(ws, s);
    hipError_t e = call.e;
    CoreOut o{coredensity, nummem, numstar, starmass, haloid, flag};
    CatShellOut so{part_ptr, part, core_ptr, cores};
    if (e == hipSuccess)
        e = catalog_watershed(nhalo, off, core_off, den, nbr, k, type, mass, peak, o, so,
                              small_max > 0 ? small_max : 131072, nshell, halo_status, ws, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(e, where);
    const int st = consume_status(ws, s, where);
    if (st != PGALF_OK) return st;
    for (int h = 0; h < nhalo; h++) {
        if (halo_status[h] == 13)
            return set_err(PGALF_ERR_CORE_OVERFLOW, "%s: halo %d's core entries outgrew its room",
                           where, h);
        if (halo_status[h])
            return set_err(PGALF_ERR_ARG, "%s: halo %d failed (%d)", where, h, halo_status[h]);
    }
    return ok();
}

// ---- launch schedule ---------------------------------------------------------------------
extern "C" int pgalf_set_schedule(const pgalf_schedule_t *bins, int nbins,
                                  int smooth_direct_max) {
    if (nbins < 0 || (nbins > 0 && !bins) || smooth_direct_max < 0 || smooth_direct_max > 64)
        return set_err(PGALF_ERR_ARG, "pgalf_set_schedule: bad arguments");
    std::vector<Schedule> t;
    for (int i = 0; i < nbins; i++) {
        const pgalf_schedule_t &b = bins[i];
        if ((i > 0 && b.nmax <= bins[i - 1].nmax) || b.knn_packed < 0 || b.knn_packed > 1 ||
            b.knn_xcd_chunk < 0 || b.knn_xcd_chunk > 1024 || b.tsc_unit_entries < 256 ||
            b.tsc_unit_entries > 65536)
            return set_err(PGALF_ERR_ARG, "pgalf_set_schedule: bad bin %d", i);
        Schedule c;
        c.nmax = b.nmax;
        c.knn_packed = b.knn_packed;
        c.knn_xcd_chunk = b.knn_xcd_chunk;
        c.tsc_unit_entries = b.tsc_unit_entries;
        t.push_back(c);
    }
    std::lock_guard<std::mutex> lk(g_sched_mu);
    g_sched = t;
    g_smooth_direct_max = smooth_direct_max;
    return ok();
}

extern "C" int pgalf_get_schedule(long long n, pgalf_schedule_t *out, int *smooth_direct_max_out) {
    const Schedule c = schedule_for(n);
    if (out) {
        out->nmax = c.nmax;
        out->knn_packed = c.knn_packed;
        out->knn_xcd_chunk = c.knn_xcd_chunk;
        out->tsc_unit_entries = c.tsc_unit_entries;
    }
    if (smooth_direct_max_out) *smooth_direct_max_out = smooth_direct_max();
    return ok();
}
