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
    unsigned *d_tile_halo = nullptr;
    long long *d_goffs = nullptr;  // smoothing groups' grid offsets, group after group
    std::vector<SmoothGroup> groups;
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
    if (b->d_tile_halo) (void)hipFree(b->d_tile_halo);
    if (b->d_goffs) (void)hipFree(b->d_goffs);
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

// The launch sequence of one stage pass (stream-ordered, no synchronisation).
hipError_t run_sequence(pgalf_batch *b, const float *x, const float *y, const float *z,
                        const float *m, float *den, float *rho_tsc, long long *nbr,
                        float *rho_knn, hipStream_t s) {
    Workspace &ws = *b->ws;
    timing_mark("batch_stage", s, false);
    timing_mark("batch_tsc", s, false);
    PGALF_HIP_TRY(tsc_deposit_batch<float>(x, y, z, m, b->bd, den, ws, s));
    timing_mark("batch_tsc", s, true);
    timing_mark("batch_smooth", s, false);
    PGALF_HIP_TRY(gaussian_smooth_batch(den, b->groups.data(), (int)b->groups.size(), s));
    timing_mark("batch_smooth", s, true);
    timing_mark("batch_findden", s, false);
    if (rho_tsc) PGALF_HIP_TRY(tsc_readback_batch<float>(x, y, z, b->bd, den, rho_tsc, ws, s));
    timing_mark("batch_findden", s, true);
    timing_mark("batch_knn", s, false);
    if (nbr || rho_knn)
        PGALF_HIP_TRY(knn_batch<float>(x, y, z, m, b->bd, b->kplan, nbr, rho_knn,
                                       rho_knn ? 1 : 0, ws, s));
    timing_mark("batch_knn", s, true);
    timing_mark("batch_stage", s, true);
    return hipGetLastError();
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
    std::vector<PChunk> chunks;
    std::vector<unsigned> tile_halo;
    b->goff.assign((size_t)nhalo + 1, 0);
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
        toff += nt;
        max_tiles = std::max(max_tiles, nt);
        for (long long c = off[h]; c < off[h + 1]; c += kChunkParticles) {
            PChunk pc;
            pc.halo = h;
            pc.b0 = c;
            pc.b1 = std::min(c + (long long)kChunkParticles, off[h + 1]);
            chunks.push_back(pc);
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
        gfirst.push_back(goffs.size());
        goffs.insert(goffs.end(), kv.second.begin(), kv.second.end());
        b->groups.push_back(grp);
    }
    hipError_t e = upload(&b->d_segs, segs);
    if (e == hipSuccess) e = upload(&b->d_chunks, chunks);
    if (e == hipSuccess) e = upload(&b->d_tile_halo, tile_halo);
    if (e == hipSuccess) e = upload(&b->d_goffs, goffs);
    for (size_t i = 0; i < b->groups.size(); i++) {
        b->groups[i].goffs = b->d_goffs + gfirst[i];
        if (e == hipSuccess) e = smooth_group_prepare(b->groups[i]);
    }
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
    b->bd.tile_halo = b->d_tile_halo;
    b->bd.ntiles = toff;
    b->bd.max_halo_tiles = max_tiles;
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
    if (!b || !x || !y || !z || !m || !den) return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    if (!use_graph) {
        if (b->gexec) {  // an eager run may resize the arena under the graph
            (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
        }
        e = b->ws->begin(s);
        if (e == hipSuccess) e = run_sequence(b, x, y, z, m, den, rho_tsc, nbr, rho_knn, s);
        return e == hipSuccess ? ok() : hip_err(e, where);
    }
    // graph mode: the sequence runs on the batch's own stream, ordered after
    // the caller's stream by events (capture is not possible on the null
    // stream, which callers commonly use)
    const std::vector<const void *> key = {x, y, z, m, den, rho_tsc, nbr, rho_knn};
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
        // size the arena with an eager pass, settle it into one block, then capture
        e = b->ws->begin(b->gstream);
        if (e == hipSuccess) e = run_sequence(b, x, y, z, m, den, rho_tsc, nbr, rho_knn, b->gstream);
        if (e == hipSuccess) e = b->ws->begin(b->gstream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->gstream);
        hipGraph_t graph = nullptr;
        if (e == hipSuccess) e = hipStreamBeginCapture(b->gstream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            timing_suspend(true);
            hipError_t er = run_sequence(b, x, y, z, m, den, rho_tsc, nbr, rho_knn, b->gstream);
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
    if (e == hipSuccess) e = hipEventRecord(b->ev_out, b->gstream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, b->ev_out, 0);
    return e == hipSuccess ? ok() : hip_err(e, where);
}

extern "C" int pgalf_batch_status(pgalf_batch_t *b, void *stream) {
    if (!b) return set_err(PGALF_ERR_ARG, "pgalf_batch_status: null batch");
    std::lock_guard<std::mutex> lk(g_api_mu);
    hipStream_t s = (hipStream_t)stream;
    if (!b->ws->status()) {
        hipError_t e = b->ws->begin(s);
        if (e != hipSuccess) return hip_err(e, "pgalf_batch_status");
    }
    return consume_status(*b->ws, s, "pgalf_batch_status");
}

extern "C" void pgalf_batch_destroy(pgalf_batch_t *b) {
    if (!b) return;
    std::lock_guard<std::mutex> lk(g_api_mu);
    (void)hipDeviceSynchronize();
    free_batch(b);
}
