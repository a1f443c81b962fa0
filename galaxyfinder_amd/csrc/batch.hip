// Batched density stage: a catalogue of halos through deposit -> smoothing
// -> read-back -> kNN lists + W4 density in one pass of launches, instead of
// one stage call per halo (the reference's MPI workers take one halo at a
// time, NewGalFinder/gfind.c:394-549, and run subhalo_den on it).
//
// The plan, built once per catalogue shape on the host, holds every
// descriptor the kernels need on the device:
//   * HaloSeg per halo (particle range, grid geometry, grid and tile offsets)
//     and the particle chunks (<= kChunkParticles, never crossing a halo) that
//     the deposit, read-back and kNN preparation kernels take one per block;
//   * the halo of every deposit tile (the deposit's work units are tiles of
//     all halos in one space, cusp tiles split across workgroups);
//   * smoothing groups: halos of one grid shape and smoothing factor share
//     FFT plans and are transformed together, one launch per axis and group;
//   * the kNN plan: per-halo box trees in one level-major array and the wave
//     table, largest halos first so the long waves start early and the small
//     halos fill the tail.
// A run issues a fixed launch sequence with no host synchronisation, no
// allocation once the batch's own arena is sized, and no memset nodes, so it
// can be captured into a HIP graph and replayed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "pgalf_internal.h"

using namespace pgalf;

struct pgalf_batch {
    int device = 0;
    int nhalo = 0, k = 0;
    std::vector<long long> goff;  // [nhalo + 1]
    BatchDev bd;
    HaloSeg *d_segs = nullptr;
    PChunk *d_chunks = nullptr;
    PChunk *d_fine = nullptr;
    unsigned *d_tile_halo = nullptr;
    long long *d_goffs = nullptr;  // smoothing groups' grid offsets, group after group
    std::vector<SmoothGroup> groups;
    SmoothBatchPlan *splan = nullptr;
    KnnBatchPlan kplan;
    Workspace *ws = nullptr;
    // graph of the launch sequence for one set of pointers
    hipStream_t gstream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<const void *> gkey;
};

namespace {

void free_batch(pgalf_batch *b) {
    if (!b) return;
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->gstream) (void)hipStreamDestroy(b->gstream);
    if (b->d_segs) (void)hipFree(b->d_segs);
    if (b->d_chunks) (void)hipFree(b->d_chunks);
    if (b->d_fine) (void)hipFree(b->d_fine);
    if (b->d_tile_halo) (void)hipFree(b->d_tile_halo);
    if (b->d_goffs) (void)hipFree(b->d_goffs);
    smooth_batch_plan_free(b->splan);
    knn_batch_plan_free(b->kplan);
    delete b->ws;
    delete b;
}

template <class T>
hipError_t upload(T **dst, const std::vector<T> &src) {
    const size_t n = src.empty() ? 1 : src.size();
    PGALF_HIP_TRY(hipMalloc(dst, sizeof(T) * n));
    if (src.empty()) return hipSuccess;
    return hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
}

// g[i] = w[i] / rho[i]: the cotangent of sum w log rho (0 where rho is 0).
__global__ void log_cotangent_k(const float *w, const float *rho, long long n, float *g) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[i] = rho[i] != 0.0f ? w[i] / rho[i] : 0.0f;
}

// The launch sequence of one stage pass (stream-ordered, no synchronisation).
hipError_t run_sequence(pgalf_batch *b, const float *x, const float *y, const float *z,
                        const float *m, float *den, float *rho_tsc, long long *nbr,
                        float *rho_knn, hipStream_t s, int rho_mode = 1) {
    Workspace &ws = *b->ws;
    timing_mark("batch_stage", s, false);
    timing_mark("batch_tsc", s, false);
    PGALF_HIP_TRY(tsc_deposit_batch<float>(x, y, z, m, b->bd, den, ws, s));
    timing_mark("batch_tsc", s, true);
    timing_mark("batch_smooth", s, false);
    PGALF_HIP_TRY(
        gaussian_smooth_batch(den, b->groups.data(), (int)b->groups.size(), b->splan, s));
    timing_mark("batch_smooth", s, true);
    // with a kNN to run, the read-back goes over its sorted slots (coherent
    // gathers, the same FP64 coordinates) and is timed inside "batch_knn"
    const bool with_knn = nbr || rho_knn;
    if (rho_tsc && !with_knn) {
        timing_mark("batch_findden", s, false);
        PGALF_HIP_TRY(tsc_readback_batch<float>(x, y, z, b->bd, den, rho_tsc, ws, s));
        timing_mark("batch_findden", s, true);
    }
    timing_mark("batch_knn", s, false);
    if (with_knn) {
        const long long *shift = b->kplan.slot_shift;
        const BatchDev &bd = b->bd;
        const SortedHook rb =
            rho_tsc ? SortedHook([&bd, den, rho_tsc, &ws, s, shift](
                                     const double *sx, const double *sy, const double *sz,
                                     const unsigned *sid, long long) {
                timing_mark("batch_findden", s, false);
                const hipError_t e =
                    tsc_readback_batch<double>(sx, sy, sz, bd, den, rho_tsc, ws, s, sid, shift);
                timing_mark("batch_findden", s, true);
                return e;
            })
                    : SortedHook();
        PGALF_HIP_TRY(knn_batch<float>(x, y, z, m, b->bd, b->kplan, nbr, rho_knn,
                                       rho_knn ? rho_mode : 0, ws, s, &rb));
    }
    timing_mark("batch_knn", s, true);
    timing_mark("batch_stage", s, true);
    return hipGetLastError();
}

struct GradArgs {
    const float *x, *y, *z, *m, *w_tsc, *w_knn;
    int rho_mode;
    float *den, *rho_tsc, *rho_knn;
    long long *nbr;
    float *gx, *gy, *gz, *gm;
    int climb_iters;
    float climb_step;
    float *cx, *cy, *cz;
    int *peak;
};

// Forward, cotangents of L = sum w_tsc log rho_tsc + sum w_knn log rho_knn,
// adjoint and peak climbing: one stream-ordered sequence.
hipError_t run_grad_sequence(pgalf_batch *b, const GradArgs &a, hipStream_t s) {
    Workspace &ws = *b->ws;
    const long long np = b->bd.np;
    PGALF_HIP_TRY(run_sequence(b, a.x, a.y, a.z, a.m, a.den, a.rho_tsc, a.nbr, a.rho_knn, s,
                               a.rho_mode));
    timing_mark("batch_bwd", s, false);
    double *acc = ws.alloc_n<double>(4 * (size_t)np);
    float *g_t = a.w_tsc ? ws.alloc_n<float>(np) : nullptr;
    float *g_k = a.w_knn ? ws.alloc_n<float>(np) : nullptr;
    if (!acc || (a.w_tsc && !g_t) || (a.w_knn && !g_k)) return hipErrorOutOfMemory;
    const unsigned nb = (unsigned)((np + 255) / 256);
    PGALF_HIP_TRY(fill_u64(reinterpret_cast<unsigned long long *>(acc), 4 * np, 0ull, s));
    if (g_t) {
        log_cotangent_k<<<nb, 256, 0, s>>>(a.w_tsc, a.rho_tsc, np, g_t);
        float *A = ws.alloc_n<float>(b->goff[b->nhalo]);
        if (!A) return hipErrorOutOfMemory;
        PGALF_HIP_TRY(tsc_deposit_batch<float>(a.x, a.y, a.z, g_t, b->bd, A, ws, s));
        PGALF_HIP_TRY(
            gaussian_smooth_batch(A, b->groups.data(), (int)b->groups.size(), b->splan, s));
        PGALF_HIP_TRY(tsc_grad_readback_batch<float>(a.x, a.y, a.z, a.m, g_t, b->bd, a.den, A,
                                                     acc, ws, s));
    }
    if (g_k) {
        log_cotangent_k<<<nb, 256, 0, s>>>(a.w_knn, a.rho_knn, np, g_k);
        PGALF_HIP_TRY(knn_batch_backward<float>(a.x, a.y, a.z, a.m, b->bd, b->kplan, a.nbr,
                                                a.rho_mode == 2 ? 1 : 0, g_k, acc, ws, s));
    }
    PGALF_HIP_TRY(grad_finalize(acc, np, a.gx, a.gy, a.gz, a.gm, s));
    timing_mark("batch_bwd", s, true);
    timing_mark("batch_climb", s, false);
    PGALF_HIP_TRY(peak_climb_batch<float>(a.x, a.y, a.z, b->bd, a.den, a.climb_iters,
                                          (double)a.climb_step, a.cx, a.cy, a.cz, a.peak, s));
    timing_mark("batch_climb", s, true);
    return hipGetLastError();
}

// Runs `seq` eagerly on the caller's stream, or (use_graph) replays the HIP
// graph captured for the pointer set `key` on the batch's own stream,
// capturing it first: an eager pass sizes the arena, the arena settles into
// one block, then the sequence is captured.  The batch stream is ordered
// after the caller's stream and back by events.
template <class Seq>
int run_maybe_graph(pgalf_batch *b, const std::vector<const void *> &key, int use_graph,
                    hipStream_t s, const char *where, Seq seq) {
    hipError_t e = hipSuccess;
    if (!use_graph) {
        if (b->gexec) {  // an eager run may resize the arena under the graph
            (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
        }
        WsCall call(*b->ws, s);
        e = call.e;
        if (e == hipSuccess) e = seq(s);
        return e == hipSuccess ? ok() : hip_err(e, where);
    }
    if (!b->gstream) {
        e = hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming);
        if (e != hipSuccess) return hip_err(e, where);
    }
    e = hipEventRecord(b->ev_in, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->gstream, b->ev_in, 0);
    if (e == hipSuccess && (!b->gexec || key != b->gkey)) {
        if (b->gexec) {
            (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
        }
        e = b->ws->begin(b->gstream);
        if (e == hipSuccess) e = seq(b->gstream);
        if (e == hipSuccess) e = b->ws->begin(b->gstream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->gstream);
        hipGraph_t graph = nullptr;
        if (e == hipSuccess) e = hipStreamBeginCapture(b->gstream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            timing_suspend(true);
            hipError_t er = seq(b->gstream);
            timing_suspend(false);
            e = hipStreamEndCapture(b->gstream, &graph);
            if (er != hipSuccess) e = er;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            b->gexec = nullptr;
            return hip_err(e, where);
        }
        b->gkey = key;
    }
    if (e == hipSuccess) e = hipGraphLaunch(b->gexec, b->gstream);
    if (e == hipSuccess) b->ws->end(b->gstream);  // the replay uses the arena
    if (e == hipSuccess) e = hipEventRecord(b->ev_out, b->gstream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, b->ev_out, 0);
    return e == hipSuccess ? ok() : hip_err(e, where);
}

}  // namespace

extern "C" pgalf_batch_t *pgalf_batch_create(int nhalo, const long long *off,
                                             const pgalf_grid_t *grids, const float *RG,
                                             int k) {
    const char *where = "pgalf_batch_create";
    if (nhalo < 1 || !off || !grids || !RG || k < 1 || k > 32 || off[0] != 0) {
        set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
        return nullptr;
    }
    for (int h = 0; h < nhalo; h++) {
        const pgalf_grid_t &g = grids[h];
        if (off[h + 1] < off[h] || g.nx < 3 || g.ny < 3 || g.nz < 3 || !(g.cellsize > 0) ||
            (long long)g.nx * g.ny * g.nz >= (1LL << 31) || !(RG[h] > 0)) {
            set_err(PGALF_ERR_ARG, "%s: bad halo %d", where, h);
            return nullptr;
        }
        if (off[h + 1] - off[h] < k) {
            set_err(PGALF_ERR_SHORT_LIST, "%s: halo %d has %lld particles < k = %d", where, h,
                    off[h + 1] - off[h], k);
            return nullptr;
        }
    }
    if (off[nhalo] >= (1LL << 32)) {
        set_err(PGALF_ERR_ARG, "%s: more than 2^32 particles", where);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_api_mu);
    pgalf_batch *b = new pgalf_batch();
    (void)hipGetDevice(&b->device);
    b->nhalo = nhalo;
    b->k = k;
    std::vector<HaloSeg> segs((size_t)nhalo);
    std::vector<PChunk> chunks, fine;
    std::vector<unsigned> tile_halo;
    b->goff.assign((size_t)nhalo + 1, 0);
    bool too_wide = false;
    int max_entries = 1;
    long long toff = 0, max_tiles = 0;
    for (int h = 0; h < nhalo; h++) {
        const pgalf_grid_t &gg = grids[h];
        HaloSeg &sg = segs[h];
        sg.g.nx = gg.nx;
        sg.g.ny = gg.ny;
        sg.g.nz = gg.nz;
        sg.g.x0 = gg.xmin;
        sg.g.y0 = gg.ymin;
        sg.g.z0 = gg.zmin;
        sg.g.h = gg.cellsize;
        sg.p0 = off[h];
        sg.np = off[h + 1] - off[h];
        sg.goff = b->goff[h];
        sg.toff = toff;
        b->goff[h + 1] = b->goff[h] + sg.g.ncells();
        const long long nt = tsc_tile_count(sg.g);
        tile_halo.insert(tile_halo.end(), (size_t)nt, (unsigned)h);
        too_wide |= !tsc_grid_fits(sg.g);
        max_entries = std::max(max_entries, tsc_tile_entries(sg.g));
        toff += nt;
        max_tiles = std::max(max_tiles, nt);
        for (long long c = off[h]; c < off[h + 1]; c += kChunkParticles) {
            PChunk pc;
            pc.halo = h;
            pc.b0 = c;
            pc.b1 = std::min(c + (long long)kChunkParticles, off[h + 1]);
            chunks.push_back(pc);
        }
        for (long long c = off[h]; c < off[h + 1]; c += kFineParticles) {
            PChunk pc;
            pc.halo = h;
            pc.b0 = c;
            pc.b1 = std::min(c + (long long)kFineParticles, off[h + 1]);
            fine.push_back(pc);
        }
    }
    // smoothing groups: one per (grid shape, smoothing factor)
    std::map<std::tuple<int, int, int, double>, std::vector<long long>> gmap;
    for (int h = 0; h < nhalo; h++) {
        const Grid &g = segs[h].g;
        gmap[std::make_tuple(g.nx, g.ny, g.nz, smooth_expfac(RG[h], g.h))].push_back(
            b->goff[h]);
    }
    std::vector<long long> goffs;
    std::vector<size_t> gfirst;  // first entry of each group in goffs
    for (auto &kv : gmap) {
        SmoothGroup grp;
        grp.nx = std::get<0>(kv.first);
        grp.ny = std::get<1>(kv.first);
        grp.nz = std::get<2>(kv.first);
        grp.expfac = std::get<3>(kv.first);
        grp.count = (int)kv.second.size();
        grp.goffs = nullptr;
        grp.hgoffs = nullptr;
        gfirst.push_back(goffs.size());
        goffs.insert(goffs.end(), kv.second.begin(), kv.second.end());
        b->groups.push_back(grp);
    }
    hipError_t e = too_wide ? hipErrorInvalidValue : hipSuccess;
    if (e == hipSuccess) e = upload(&b->d_segs, segs);
    if (e == hipSuccess) e = upload(&b->d_chunks, chunks);
    if (e == hipSuccess) e = upload(&b->d_fine, fine);
    if (e == hipSuccess) e = upload(&b->d_tile_halo, tile_halo);
    if (e == hipSuccess) e = upload(&b->d_goffs, goffs);
    for (size_t i = 0; i < b->groups.size(); i++) {
        b->groups[i].goffs = b->d_goffs + gfirst[i];
        b->groups[i].hgoffs = goffs.data() + gfirst[i];
        if (e == hipSuccess) e = smooth_group_prepare(b->groups[i]);
    }
    if (e == hipSuccess) b->splan = smooth_batch_plan_create(b->groups.data(),
                                                             (int)b->groups.size(), e);
    for (auto &grp : b->groups) grp.hgoffs = nullptr;
    if (e == hipSuccess) e = knn_batch_plan(off, nhalo, k, b->kplan);
    if (e != hipSuccess) {
        hip_err(e, where);
        free_batch(b);
        return nullptr;
    }
    b->bd.nhalo = nhalo;
    b->bd.np = off[nhalo];
    b->bd.segs = b->d_segs;
    b->bd.chunks = b->d_chunks;
    b->bd.nchunks = (long long)chunks.size();
    b->bd.fine = b->d_fine;
    b->bd.nfine = (long long)fine.size();
    b->bd.tile_halo = b->d_tile_halo;
    b->bd.ntiles = toff;
    b->bd.max_halo_tiles = max_tiles;
    b->bd.max_tile_entries = max_entries;
    b->bd.sched = b->kplan.sched;  // the schedule of the batch's total, fixed here
    b->ws = new Workspace(b->device);
    ok();
    return b;
}

extern "C" long long pgalf_batch_grid_offset(const pgalf_batch_t *b, int h) {
    if (!b || h < 0 || h > b->nhalo) return -1;
    return b->goff[(size_t)h];
}

extern "C" int pgalf_batch_run(pgalf_batch_t *b, const float *x, const float *y,
                               const float *z, const float *m, float *den, float *rho_tsc,
                               long long *nbr, float *rho_knn, int use_graph, void *stream) {
    const char *where = "pgalf_batch_run";
    if (!b || !x || !y || !z || !m || !den || !aligned8(den))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    const std::vector<const void *> key = {x, y, z, m, den, rho_tsc, nbr, rho_knn};
    return run_maybe_graph(b, key, use_graph, (hipStream_t)stream, where, [&](hipStream_t s) {
        return run_sequence(b, x, y, z, m, den, rho_tsc, nbr, rho_knn, s);
    });
}

extern "C" int pgalf_batch_run_grad(pgalf_batch_t *b, const float *x, const float *y,
                                    const float *z, const float *m, const float *w_tsc,
                                    const float *w_knn, int rho_mode, float *den,
                                    float *rho_tsc, long long *nbr, float *rho_knn, float *gx,
                                    float *gy, float *gz, float *gm, int climb_iters,
                                    float climb_step, float *cx, float *cy, float *cz,
                                    int *peak, int use_graph, void *stream) {
    const char *where = "pgalf_batch_run_grad";
    if (!b || !x || !y || !z || !m || !den || !aligned8(den) || !rho_tsc || !nbr || !rho_knn ||
        !gx || !gy ||
        !gz || !gm || !cx || !cy || !cz || !peak || (rho_mode != 1 && rho_mode != 2) ||
        climb_iters < 0 || !(climb_step >= 0.0f))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    GradArgs a{x, y, z, m, w_tsc, w_knn, rho_mode, den, rho_tsc, rho_knn, nbr,
               gx, gy, gz, gm, climb_iters, climb_step, cx, cy, cz, peak};
    // the iteration count and step are baked into a captured graph: part of
    // the key (the step by its bit pattern, so any two distinct steps differ)
    unsigned step_bits = 0;
    memcpy(&step_bits, &climb_step, sizeof(step_bits));
    const std::vector<const void *> key = {
        x, y, z, m, w_tsc, w_knn, den, rho_tsc, nbr, rho_knn, gx, gy, gz, gm, cx, cy, cz, peak,
        reinterpret_cast<const void *>((long long)rho_mode),
        reinterpret_cast<const void *>((long long)climb_iters),
        reinterpret_cast<const void *>((unsigned long long)step_bits)};
    return run_maybe_graph(b, key, use_graph, (hipStream_t)stream, where,
                           [&](hipStream_t s) { return run_grad_sequence(b, a, s); });
}

extern "C" int pgalf_batch_status(pgalf_batch_t *b, void *stream) {
    if (!b) return set_err(PGALF_ERR_ARG, "pgalf_batch_status: null batch");
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    if (!b->ws->status()) {
        WsCall call(*b->ws, s);
        if (call.e != hipSuccess) return hip_err(call.e, "pgalf_batch_status");
    }
    return consume_status(*b->ws, s, "pgalf_batch_status");
}

extern "C" long long pgalf_batch_workspace_bytes(const pgalf_batch_t *b, int reserved) {
    if (!b || !b->ws) return -1;
    std::lock_guard<std::mutex> lk(g_api_mu);
    return (long long)(reserved ? b->ws->reserved() : b->ws->high_water());
}

extern "C" void pgalf_batch_destroy(pgalf_batch_t *b) {
    if (!b) return;
    std::lock_guard<std::mutex> lk(g_api_mu);
    (void)hipDeviceSynchronize();
    free_batch(b);
}
