// FoF catalogue loader and GALFIND writer (host code of the library).
//
// Loader (gfind.c:25-91,221-241,363-371; subhaloden.mod6.c:79-177,
// 2810-2837): the halo catalogue file holds a 7-float header and one HaloQ
// record per halo, the member file each halo's DM, gas, sink and star
// records in that order.  A batch of halos is chosen from the HaloQ records
// (smaller halos are skipped over); host threads then read their halos'
// records with positional reads and localise each halo as subhalo_den does
// before the density stage: Dump2BP subtracts the per-axis minimum
// (FP64), mklocalize moves the particles below the widest gap of a halo
// that spans the periodic box by one box size.  The reference's gap search
// sorts the coordinates with a comparator that compares the element
// addresses (subhaloden.mod6.c:57-64); under glibc's merge sort that leaves
// them in input order, so the gap it finds is the widest step between
// consecutive particles of the file.  PGALF_UNWRAP_REFERENCE reproduces
// that (the default: results identical to the reference build);
// PGALF_UNWRAP_SORTED sorts the values, which is what the code intends.
// Each halo also gets the TSC grid lagFindStellarCore would size
// (nnost.c:1040-1058) and the link length of FREAD (gfind.c:51).
//
// Writer (gfind.c:557-732 write_data): per halo with subhalos, a HaloInfo
// record to GALFIND.DATA and GALCATALOG.LIST, then per subhalo a SubInfo to
// both and its member records (DM, gas, sink, star, each in particle order)
// to GALFIND.DATA; the particles of no subhalo go to the background file.
// Sums run in FP64 in particle order, as the reference's.
#include <pthread.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <system_error>
#include <thread>
#include <vector>

#include "pgalf_catalog.h"
#include "pgalf_internal.h"

// status of the calling thread (capi.hip)
int set_err(int code, const char *fmt, ...);
int ok();

namespace {

// Record layouts of ramses.h:66-159 / tree.h:55-60 / defs.h:3-27 under the
// reference Makefile's flags (NCHEM = 3, no NBODY), checked against the
// reference build by the tests.
constexpr size_t kPmBytes = 136, kGasBytes = 112, kSinkBytes = 168;
constexpr int kTypeStar = 1, kTypeDm = 2, kTypeGas = 3, kTypeSink = 4;
constexpr int kNotMember = -9999;
static_assert(sizeof(pgalf_haloq_t) == 96, "HaloQ record");
static_assert(sizeof(pgalf_halo_info_t) == 112 && sizeof(pgalf_sub_info_t) == 112,
              "HaloInfo / SubInfo records");

size_t record_bytes(int type) {
    return type == kTypeGas ? kGasBytes : (type == kTypeSink ? kSinkBytes : kPmBytes);
}

double rd_d(const unsigned char *r, size_t off) {
    double v;
    std::memcpy(&v, r + off, sizeof(v));
    return v;
}
float rd_f(const unsigned char *r, size_t off) {
    float v;
    std::memcpy(&v, r + off, sizeof(v));
    return v;
}

// Mass and velocity of a record as FREAD / GetHaloInfo read them.
void record_mv(const unsigned char *r, int type, double &m, double &vx, double &vy, double &vz) {
    if (type == kTypeGas) {
        m = (double)rd_f(r, 76);
        vx = (double)rd_f(r, 32);
        vy = (double)rd_f(r, 36);
        vz = (double)rd_f(r, 40);
    } else {
        m = rd_d(r, 48);
        vx = rd_d(r, 24);
        vy = rd_d(r, 32);
        vz = rd_d(r, 40);
    }
}

// A host array that grows without zero-filling: the loader's threads write
// every element, so the pages are first touched by those threads.
template <class T>
struct HostArray {
    std::unique_ptr<T[]> p;
    size_t n = 0, cap = 0;
    bool resize(size_t m) {
        if (m > cap) {
            p.reset(new (std::nothrow) T[m]);
            cap = p ? m : 0;
        }
        n = p ? m : 0;
        return p || m == 0;
    }
    T *data() { return p.get(); }
    const T *data() const { return p.get(); }
    T &operator[](size_t i) { return p[i]; }
    const T &operator[](size_t i) const { return p[i]; }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
};

// Host threads for the loader: at most 16, one if the machine has one.
unsigned loader_threads() {
    return std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

}  // namespace

struct pgalf_catalog {
    FILE *hf = nullptr, *mf = nullptr;
    int unwrap = PGALF_UNWRAP_REFERENCE;
    pgalf_cat_header_t hdr{};
    bool have_pending = false;  // a HaloQ read but not yet consumed
    pgalf_haloq_t pending{};
    // the last batch
    std::vector<pgalf_haloq_t> haloq;
    std::vector<long long> off, roff;  // particle and record-byte offsets per halo
    std::unique_ptr<unsigned char[]> rec;  // records of the batch (not zero-filled)
    size_t rec_bytes = 0, rec_cap = 0;
    HostArray<double> x, y, z, m, vx, vy, vz;
    HostArray<int> type;
    HostArray<float> link02;
    std::vector<double> origin;
    std::vector<pgalf_grid_t> grids;
};

namespace {

// mklocalize along one axis (subhaloden.mod6.c:79-177), on the coordinates
// after Dump2BP.
void unwrap_axis(double *c, long long n, float size, int mode) {
    const int nblur = 3;  // defs.h:31
    float lo = size, hi = -size;
    for (long long i = 0; i < n; i++) {
        lo = (float)(lo < c[i] ? (double)lo : c[i]);
        hi = (float)(hi > c[i] ? (double)hi : c[i]);
    }
    if (!(std::floor(lo) <= 0 + nblur && std::ceil(hi) >= size - nblur) || n < 2) return;
    std::vector<float> t(n);
    for (long long i = 0; i < n; i++) t[i] = (float)c[i];
    if (mode == PGALF_UNWRAP_SORTED) std::sort(t.begin(), t.end());
    float maxdist = -10.0f;
    long long up = 0;
    for (long long i = 0; i < n - 1; i++) {
        const float d = t[i + 1] - t[i];
        if (d > maxdist) {
            up = i;
            maxdist = d;
        }
    }
    // the reference compares the FP64 coordinate with the float cut, so a
    // particle just above its float rounding stays behind; the sorted mode
    // compares the float values the gap was found among
    const float cut = t[up];
    for (long long i = 0; i < n; i++)
        if (mode == PGALF_UNWRAP_SORTED ? (float)c[i] <= cut : c[i] <= (double)cut)
            c[i] += (double)size;
}

// Dump2BP (subhaloden.mod6.c:2810-2837) and mklocalize of one halo; the
// grid of nnost.c:1040-1058.
void localise(pgalf_catalog *c, long long h) {
    const long long a = c->off[h], b = c->off[h + 1], n = b - a;
    double *ax[3] = {c->x.data() + a, c->y.data() + a, c->z.data() + a};
    for (int d = 0; d < 3; d++) {
        double mn = 1.0e30;
        for (long long i = 0; i < n; i++) mn = ax[d][i] < mn ? ax[d][i] : mn;
        c->origin[3 * h + d] = mn;
        for (long long i = 0; i < n; i++) ax[d][i] = ax[d][i] - mn;
    }
    for (int d = 0; d < 3; d++) unwrap_axis(ax[d], n, c->hdr.size, c->unwrap);
    int nn[3];
    double lo[3];
    pgalf::lag_grid_box(ax[0], ax[1], ax[2], n, 1, nn, lo);
    pgalf_grid_t &g = c->grids[h];
    g.nx = nn[0];
    g.ny = nn[1];
    g.nz = nn[2];
    g.xmin = lo[0];
    g.ymin = lo[1];
    g.zmin = lo[2];
    g.cellsize = pgalf::kLagCellSize;
}

}  // namespace

extern "C" pgalf_catalog_t *pgalf_catalog_open(const char *halo_cat, const char *members,
                                               int unwrap) {
    if (!halo_cat || !members ||
        (unwrap != PGALF_UNWRAP_REFERENCE && unwrap != PGALF_UNWRAP_SORTED)) {
        set_err(PGALF_ERR_ARG, "pgalf_catalog_open: bad arguments");
        return nullptr;
    }
    pgalf_catalog *c = new pgalf_catalog;
    c->unwrap = unwrap;
    c->hf = std::fopen(halo_cat, "rb");
    c->mf = std::fopen(members, "rb");
    float h[7];
    if (!c->hf || !c->mf || std::fread(h, sizeof(float), 7, c->hf) != 7) {
        set_err(PGALF_ERR_ARG, "pgalf_catalog_open: cannot read %s / %s", halo_cat, members);
        pgalf_catalog_close(c);
        return nullptr;
    }
    c->hdr.size = h[0];
    c->hdr.hubble = h[1];
    c->hdr.omep = h[2];
    c->hdr.omepb = h[3];
    c->hdr.omeplam = h[4];
    c->hdr.amax = h[5];
    c->hdr.anow = h[6];
    ok();
    return c;
}

extern "C" int pgalf_catalog_header(const pgalf_catalog_t *c, pgalf_cat_header_t *out) {
    if (!c || !out) return set_err(PGALF_ERR_ARG, "pgalf_catalog_header: bad arguments");
    *out = c->hdr;
    return ok();
}

extern "C" long long pgalf_catalog_next(pgalf_catalog_t *c, long long max_halos,
                                        long long max_particles, int min_np, long long *np_out) {
    const char *where = "pgalf_catalog_next";
    if (!c || max_halos < 1 || max_particles < 1) {
        set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
        return -1;
    }
    c->haloq.clear();
    c->off.assign(1, 0);
    c->roff.assign(1, 0);
    long long tot = 0;
    size_t rbytes = 0;
    // choose the batch's halos from their HaloQ records; a halo below
    // min_np is passed by (gfind.c:388 reads it and hands it to no one):
    // skip[h] bytes of the member file lie before chosen halo h
    std::vector<size_t> skip;
    size_t gap = 0;
    while ((long long)c->haloq.size() < max_halos) {
        pgalf_haloq_t q;
        if (c->have_pending) {
            q = c->pending;
            c->have_pending = false;
        } else if (std::fread(&q, sizeof(q), 1, c->hf) != 1) {
            break;
        }
        const size_t qb = (size_t)(q.npdm + q.npstar) * kPmBytes + (size_t)q.npgas * kGasBytes +
                          (size_t)q.npsink * kSinkBytes;
        if (q.npdm + q.npgas + q.npsink + q.npstar != q.np) {
            set_err(PGALF_ERR_ARG, "%s: HaloQ type counts do not add up to np", where);
            return -1;
        }
        if ((long long)q.np < min_np) {
            gap += qb;
            continue;
        }
        if (!c->haloq.empty() && tot + (long long)q.np > max_particles) {
            c->pending = q;
            c->have_pending = true;
            break;
        }
        c->haloq.push_back(q);
        skip.push_back(gap);
        gap = 0;
        tot += (long long)q.np;
        rbytes += qb;
        c->off.push_back(tot);
        c->roff.push_back((long long)rbytes);
    }
    if (rbytes > c->rec_cap) {
        c->rec.reset(new (std::nothrow) unsigned char[rbytes]);
        c->rec_cap = c->rec ? rbytes : 0;
        if (!c->rec) {
            set_err(PGALF_ERR_NOMEM, "%s: host allocation failed", where);
            return -1;
        }
    }
    c->rec_bytes = rbytes;
    // member-file position of each chosen halo's records
    const long long nh = (long long)c->haloq.size();
    const off_t base = ftello(c->mf);
    if (base < 0) {
        set_err(PGALF_ERR_ARG, "%s: cannot tell the member file's position", where);
        return -1;
    }
    std::vector<long long> fpos(nh);
    long long at = (long long)base;
    for (long long h = 0; h < nh; h++) {
        at += (long long)skip[h];
        fpos[h] = at;
        at += c->roff[h + 1] - c->roff[h];
    }
    const long long end = at + (long long)gap;  // after the small halos that follow
    bool grown = true;
    for (auto *v : {&c->x, &c->y, &c->z, &c->m, &c->vx, &c->vy, &c->vz}) grown &= v->resize(tot);
    grown = grown && c->type.resize(tot) && c->link02.resize(tot);
    if (!grown) {
        set_err(PGALF_ERR_NOMEM, "%s: host allocation failed", where);
        return -1;
    }
    c->origin.resize(3 * (size_t)nh);
    c->grids.resize(nh);
    const double omep = (double)c->hdr.omep;
    const int fd = fileno(c->mf);
    std::atomic<bool> short_file(false);
    // halos [h0, h1): one positional read per run of consecutive records,
    // then the particles' fields, link length and localisation
    auto unpack = [&](long long h0, long long h1) {
        for (long long h = h0; h < h1;) {
            long long e = h + 1;
            while (e < h1 && skip[e] == 0) e++;
            unsigned char *dst = c->rec.get() + c->roff[h];
            size_t left = (size_t)(c->roff[e] - c->roff[h]);
            off_t pos = (off_t)fpos[h];
            while (left) {
                const ssize_t got = pread(fd, dst, left, pos);
                if (got <= 0) {
                    short_file = true;
                    return;
                }
                dst += got;
                pos += got;
                left -= (size_t)got;
            }
            h = e;
        }
        // FREAD's link length of the last mass seen (consecutive particles
        // mostly share one mass): the same value the formula gives
        double last_m = 0.0;
        float last_l = 0.0f;
        bool have = false;
        for (long long h = h0; h < h1; h++) {
            const pgalf_haloq_t &q = c->haloq[h];
            const unsigned char *r = c->rec.get() + c->roff[h];
            long long p = c->off[h];
            const struct {
                unsigned long long n;
                int type;
            } parts[4] = {{q.npdm, kTypeDm}, {q.npgas, kTypeGas}, {q.npsink, kTypeSink},
                          {q.npstar, kTypeStar}};
            for (const auto &pt : parts) {
                const size_t rb = record_bytes(pt.type);
                for (unsigned long long i = 0; i < pt.n; i++, p++, r += rb) {
                    c->x[p] = rd_d(r, 0);
                    c->y[p] = rd_d(r, 8);
                    c->z[p] = rd_d(r, 16);
                    record_mv(r, pt.type, c->m[p], c->vx[p], c->vy[p], c->vz[p]);
                    c->type[p] = pt.type;
                    if (!have || !(c->m[p] == last_m)) {
                        // gfind.c:51: the quotient in long double, C's pow (double)
                        const double q3 = (double)((long double)c->m[p] / 2.7755e11L /
                                                   (long double)omep);
                        last_l = (float)(0.2 * std::pow(q3, (double)0.33333333333333333333L));
                        last_m = c->m[p];
                        have = true;
                    }
                    c->link02[p] = last_l;
                }
            }
            localise(c, h);
        }
    };
    // halos split over host threads by particle count
    const unsigned nt = loader_threads();
    if (nh > 1 && tot > 200000 && nt > 1) {
        std::vector<std::thread> th;
        long long h = 0;
        for (unsigned t = 0; t < nt && h < nh; t++) {
            const long long target = tot * (long long)(t + 1) / nt;
            long long e = h;
            while (e < nh && (c->off[e + 1] <= target || e == h)) e++;
            if (t + 1 == nt) e = nh;
            try {
                th.emplace_back(unpack, h, e);
            } catch (const std::system_error &) {
                unpack(h, e);
            }
            h = e;
        }
        if (h < nh) unpack(h, nh);
        for (auto &t : th) t.join();
    } else {
        unpack(0, nh);
    }
    if (short_file || fseeko(c->mf, (off_t)end, SEEK_SET) != 0) {
        set_err(PGALF_ERR_ARG, "%s: member file shorter than the catalogue", where);
        return -1;
    }
    if (np_out) *np_out = tot;
    ok();
    return nh;
}

extern "C" int pgalf_catalog_fetch(const pgalf_catalog_t *c, double *x, double *y, double *z,
                                   double *mass, int *type, float *link02, double *vx,
                                   double *vy, double *vz, long long *offsets,
                                   long long *record_offsets, pgalf_haloq_t *haloq,
                                   double *origin, pgalf_grid_t *grids) {
    if (!c) return set_err(PGALF_ERR_ARG, "pgalf_catalog_fetch: bad arguments");
    // the copies, split over host threads by bytes
    struct Copy {
        void *dst;
        const void *src;
        size_t bytes;
    };
    std::vector<Copy> jobs;
    auto cp = [&jobs](auto *dst, const auto &src) {
        if (dst && !src.empty()) jobs.push_back({dst, src.data(), sizeof(src[0]) * src.size()});
    };
    cp(x, c->x);
    cp(y, c->y);
    cp(z, c->z);
    cp(mass, c->m);
    cp(type, c->type);
    cp(link02, c->link02);
    cp(vx, c->vx);
    cp(vy, c->vy);
    cp(vz, c->vz);
    cp(offsets, c->off);
    cp(record_offsets, c->roff);
    cp(haloq, c->haloq);
    cp(origin, c->origin);
    cp(grids, c->grids);
    size_t total = 0;
    for (const Copy &j : jobs) total += j.bytes;
    // bytes [lo, hi) of the jobs laid end to end
    auto run = [&jobs](size_t lo, size_t hi) {
        size_t at = 0;
        for (const Copy &j : jobs) {
            const size_t a = std::max(lo, at), b = std::min(hi, at + j.bytes);
            if (a < b)
                std::memcpy((char *)j.dst + (a - at), (const char *)j.src + (a - at), b - a);
            at += j.bytes;
        }
    };
    const unsigned nt = total >= ((size_t)8 << 20) ? loader_threads() : 1u;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) {
        const size_t lo = total * t / nt, hi = total * (t + 1) / nt;
        try {
            th.emplace_back(run, lo, hi);
        } catch (const std::system_error &) {
            run(lo, hi);
        }
    }
    run(0, total / nt);
    for (auto &t : th) t.join();
    return ok();
}

extern "C" const void *pgalf_catalog_records(const pgalf_catalog_t *c, long long *bytes) {
    if (!c) return nullptr;
    if (bytes) *bytes = (long long)c->rec_bytes;
    return c->rec.get();
}

extern "C" void pgalf_catalog_close(pgalf_catalog_t *c) {
    if (!c) return;
    if (c->hf) std::fclose(c->hf);
    if (c->mf) std::fclose(c->mf);
    delete c;
}

// ---- writer -------------------------------------------------------------------------
struct pgalf_writer {
    FILE *data = nullptr, *list = nullptr, *bg = nullptr;
};

extern "C" pgalf_writer_t *pgalf_writer_open(const char *galfind_data, const char *catalog_list,
                                             const char *background) {
    if (!galfind_data || !catalog_list) {
        set_err(PGALF_ERR_ARG, "pgalf_writer_open: bad arguments");
        return nullptr;
    }
    pgalf_writer *w = new pgalf_writer;
    w->data = std::fopen(galfind_data, "wb");
    w->list = std::fopen(catalog_list, "wb");
    if (background) w->bg = std::fopen(background, "wb");
    if (!w->data || !w->list || (background && !w->bg)) {
        set_err(PGALF_ERR_ARG, "pgalf_writer_open: cannot create the output files");
        pgalf_writer_close(w);
        return nullptr;
    }
    ok();
    return w;
}

namespace {

// GetSubHaloInfo (gfind.c:618-686) of the particles with p2halo == id, and
// their records grouped DM, gas, sink, star.
void sub_info(int np, const int *type, const unsigned char *const *rp, const int *p2halo, int id,
              const double *px, const double *py, const double *pz, pgalf_sub_info_t &s,
              std::vector<unsigned char> &out) {
    std::memset(&s, 0, sizeof(s));
    std::vector<int> idx[4];  // dm, gas, sink, star
    for (int i = 0; i < np; i++) {
        if (p2halo[i] != id) continue;
        double m, vx, vy, vz;
        record_mv(rp[i], type[i], m, vx, vy, vz);
        switch (type[i]) {
        case kTypeDm: idx[0].push_back(i); s.npdm++; s.mdm += m; break;
        case kTypeGas: idx[1].push_back(i); s.npgas++; s.mgas += m; break;
        case kTypeSink: idx[2].push_back(i); s.npsink++; s.msink += m; break;
        case kTypeStar: idx[3].push_back(i); s.npstar++; s.mstar += m; break;
        default: break;
        }
        s.x += m * px[i];
        s.y += m * py[i];
        s.z += m * pz[i];
        s.vx += m * vx;
        s.vy += m * vy;
        s.vz += m * vz;
        s.totm += m;
        s.npall++;
    }
    s.x /= s.totm;
    s.y /= s.totm;
    s.z /= s.totm;
    s.vx /= s.totm;
    s.vy /= s.totm;
    s.vz /= s.totm;
    out.clear();
    for (const auto &v : idx)
        for (int i : v) out.insert(out.end(), rp[i], rp[i] + record_bytes(type[i]));
}

}  // namespace

extern "C" int pgalf_writer_halo(pgalf_writer_t *w, int np, const int *type, const void *records,
                                 const int *p2halo, int mpeak) {
    const char *where = "pgalf_writer_halo";
    if (!w || np < 0 || (np > 0 && (!type || !records || !p2halo)))
        return set_err(PGALF_ERR_ARG, "%s: bad arguments", where);
    if (mpeak <= 0) return ok();
    std::vector<const unsigned char *> rp(np);
    std::vector<double> px(np), py(np), pz(np);
    const unsigned char *r = static_cast<const unsigned char *>(records);
    for (int i = 0; i < np; i++) {
        if (type[i] < kTypeStar || type[i] > kTypeSink)
            return set_err(PGALF_ERR_ARG, "%s: particle %d has type %d", where, i, type[i]);
        rp[i] = r;
        px[i] = rd_d(r, 0);
        py[i] = rd_d(r, 8);
        pz[i] = rd_d(r, 16);
        r += record_bytes(type[i]);
    }
    // GetHaloInfo (gfind.c:557-617)
    pgalf_halo_info_t hi;
    std::memset(&hi, 0, sizeof(hi));
    for (int i = 0; i < np; i++) {
        double m, vx, vy, vz;
        record_mv(rp[i], type[i], m, vx, vy, vz);
        switch (type[i]) {
        case kTypeDm: hi.ndm++; hi.mdm += m; break;
        case kTypeSink: hi.nsink++; hi.msink += m; break;
        case kTypeStar: hi.nstar++; hi.mstar += m; break;
        case kTypeGas: hi.ngas++; hi.mgas += m; break;
        default: break;
        }
        hi.totm += m;
        hi.x += m * px[i];
        hi.y += m * py[i];
        hi.z += m * pz[i];
        hi.vx += m * vx;
        hi.vy += m * vy;
        hi.vz += m * vz;
    }
    hi.npall = np;
    hi.x /= hi.totm;
    hi.y /= hi.totm;
    hi.z /= hi.totm;
    hi.vx /= hi.totm;
    hi.vy /= hi.totm;
    hi.vz /= hi.totm;
    hi.nsub = mpeak;
    bool okw = std::fwrite(&hi, sizeof(hi), 1, w->data) == 1 &&
               std::fwrite(&hi, sizeof(hi), 1, w->list) == 1;
    std::vector<unsigned char> buf;
    for (int s = 0; s < mpeak && okw; s++) {
        pgalf_sub_info_t si;
        sub_info(np, type, rp.data(), p2halo, s, px.data(), py.data(), pz.data(), si, buf);
        okw = std::fwrite(&si, sizeof(si), 1, w->data) == 1 &&
              std::fwrite(&si, sizeof(si), 1, w->list) == 1 &&
              (buf.empty() || std::fwrite(buf.data(), 1, buf.size(), w->data) == buf.size());
    }
    if (okw && w->bg) {
        pgalf_sub_info_t si;
        sub_info(np, type, rp.data(), p2halo, kNotMember, px.data(), py.data(), pz.data(), si,
                 buf);
        okw = std::fwrite(&si, sizeof(si), 1, w->bg) == 1 &&
              (buf.empty() || std::fwrite(buf.data(), 1, buf.size(), w->bg) == buf.size());
    }
    if (!okw) return set_err(PGALF_ERR_ARG, "%s: write failed", where);
    return ok();
}

extern "C" int pgalf_writer_close(pgalf_writer_t *w) {
    if (!w) return ok();
    bool good = true;
    for (FILE *f : {w->data, w->list, w->bg})
        if (f && std::fclose(f) != 0) good = false;
    delete w;
    return good ? ok() : set_err(PGALF_ERR_ARG, "pgalf_writer_close: close failed");
}
