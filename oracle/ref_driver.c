/*
 * ref_driver.c -- thin C driver around the REFERENCE's own compiled sources
 * (NewGalFinder/tsc_omp2.c, nnost.c, Memory2.c, built in place from
 * /root/reference by oracle/Makefile into oracle/_ref/libpgalf_ref.so).
 *
 * TEST INFRASTRUCTURE ONLY: used to pin the oracle (tests/) and to write the
 * golden fixtures (tests/golden/make_golden.py).  It packs plain SoA arrays
 * into the reference's particle records and calls the reference functions;
 * the loops that mirror lagFindStellarCore's neighbour-list driver are this
 * file's own code calling the reference's Find_Near_PreCopied.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#include "ramses.h"
#include "tree.h"
#include "params.h"
#include "Memory.h"

/* Reference entry points (tsc_omp2.c:40,166; nnost.c:199,330,375,536,560). */
void assign_density_TSC(SimpleBasicParticleType *, int, float *, int, int, int,
                        double, double, double, double, int);
void findDen(SimpleBasicParticleType *, int, float *, double, double, double,
             double, float *, int, int, int);
void findsphdensity(SimpleBasicParticleType *, int, long long *, int, float *);
void new_Make_Tree_Near(TStruct *, size_t, TPtlStruct *, size_t, int);
float W4(float, float);

/* Record layout of nnost.c:273-277 (file-local there, restated here). */
typedef struct {
    dptype dist2;
    float mass;
    long long indx;
} ref_neighbor;
int Find_Near_PreCopied(particle *, int, TStruct *, TPtlStruct *, ref_neighbor *, int);
int Find_Near(particle *, int, TStruct *, TPtlStruct *, float *, long long *,
              dptype *, float *);

static int ref_pool_ready = 0;

int ref_init(long long bytes) {
    if (ref_pool_ready) return 1;
    ref_pool_ready = (int)Make_Total_Memory((INT8)bytes);
    return ref_pool_ready;
}

static SimpleBasicParticleType *ref_pack(const double *x, const double *y,
                                         const double *z, const double *m,
                                         const int *type, long long np) {
    SimpleBasicParticleType *bp =
        (SimpleBasicParticleType *)calloc((size_t)np, sizeof(SimpleBasicParticleType));
    for (long long i = 0; i < np; i++) {
        bp[i].type = type ? (unsigned int)type[i] : 1u;
        bp[i].mass = m ? m[i] : 1.0;
        bp[i].x = x[i];
        bp[i].y = y[i];
        bp[i].z = z[i];
        bp[i].indx = (idtype)i;
        bp[i].bp = NULL;
    }
    return bp;
}

int ref_sizeof_bp(void) { return (int)sizeof(SimpleBasicParticleType); }

void ref_tsc(const double *x, const double *y, const double *z, const double *m,
             const int *type, long long np, float *den, int nx, int ny, int nz,
             double xmin, double ymin, double zmin, double cellsize, int ptype,
             int nthreads) {
    ref_init(4LL << 30);
    SimpleBasicParticleType *bp = ref_pack(x, y, z, m, type, np);
    omp_set_num_threads(nthreads);
    assign_density_TSC(bp, (int)np, den, nx, ny, nz, xmin, ymin, zmin, cellsize, ptype);
    free(bp);
}

void ref_findden(const double *x, const double *y, const double *z, long long np,
                 float *densph, double xmin, double ymin, double zmin,
                 double cellsize, float *den, int nx, int ny, int nz, int nthreads) {
    SimpleBasicParticleType *bp = ref_pack(x, y, z, NULL, NULL, np);
    omp_set_num_threads(nthreads);
    findDen(bp, (int)np, densph, xmin, ymin, zmin, cellsize, den, nx, ny, nz);
    free(bp);
}

float ref_W4(float dist, float h) { return W4(dist, h); }

/* Tree set-up as lagFindStellarCore does it (nnost.c:1298-1329): unit masses,
 * serialized build, MAX(np/2, 650000) nodes. */
static TStruct *ref_tree(const double *x, const double *y, const double *z,
                         long long np, TPtlStruct **ptl_out) {
    TPtlStruct *ptl = (TPtlStruct *)calloc((size_t)np, sizeof(TPtlStruct));
    for (long long i = 0; i < np; i++) {
        ptl[i].type = TYPE_PTL;
        ptl[i].x = x[i];
        ptl[i].y = y[i];
        ptl[i].z = z[i];
        ptl[i].sibling = ptl + (i + 1);
        ptl[i].mass = 1;
    }
    ptl[np - 1].sibling = NULL;
    size_t nnode = (size_t)(np / 2 > 650000 ? np / 2 : 650000);
    TStruct *tree = (TStruct *)calloc(nnode, sizeof(TStruct));
    new_Make_Tree_Near(tree, nnode, ptl, (size_t)np, SERIALIZED);
    *ptl_out = ptl;
    return tree;
}

/* Cold-start Find_Near_PreCopied per particle (exact kNN, first-seen ties).
 * Returns the number of queries that came back short. */
long long ref_knn_precopied(const double *x, const double *y, const double *z,
                            long long np, int k, long long *idx, double *d2,
                            int nthreads) {
    TPtlStruct *ptl;
    TStruct *tree = ref_tree(x, y, z, np, &ptl);
    long long shortq = 0;
    omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : shortq)
    {
        ref_neighbor nb[MAX_NUM_NEAR];
#pragma omp for schedule(dynamic, 64)
        for (long long i = 0; i < np; i++) {
            particle pp;
            pp.x = x[i]; pp.y = y[i]; pp.z = z[i];
            int n = Find_Near_PreCopied(&pp, k, tree, ptl, nb, 0);
            if (n != k) shortq++;
            for (int j = 0; j < k; j++) {
                idx[i * k + j] = j < n ? nb[j].indx : -1;
                if (d2) d2[i * k + j] = j < n ? nb[j].dist2 : -1.0;
            }
        }
    }
    free(tree);
    free(ptl);
    return shortq;
}

static int ref_cmp_nb(const void *a, const void *b) {
    const ref_neighbor *aa = (const ref_neighbor *)a, *bb = (const ref_neighbor *)b;
    if (aa->dist2 < bb->dist2) return -1;
    if (aa->dist2 > bb->dist2) return 1;
    return 0;
}

/* Warm-started lists on the TSC node cells, the np >= 1e5 branch of
 * lagFindStellarCore (nnost.c:1361-1461, shallow cells only, one thread per
 * cell): previous list re-measured, qsort-ed, then Find_Near_PreCopied.
 * Cells with DEEPSIZE or more members are handled with the same warm start
 * serially (their OpenMP split changes only the schedule).  Used as the
 * reference-kind CPU timing and to confirm warm == cold on the fixtures. */
long long ref_knn_warm(const double *x, const double *y, const double *z,
                       long long np, int k, double xmin, double ymin, double zmin,
                       double cellsize, int nx, int ny, int nz, long long *idx,
                       int nthreads) {
    TPtlStruct *ptl;
    TStruct *tree = ref_tree(x, y, z, np, &ptl);
    long long mx = 2 * (nx / 2 + 1);
    long long ncells = mx * (long long)ny * nz;
    long long *head = (long long *)malloc(sizeof(long long) * (size_t)ncells);
    long long *next = (long long *)malloc(sizeof(long long) * (size_t)np);
    for (long long c = 0; c < ncells; c++) head[c] = -1;
    for (long long i = 0; i < np; i++) {
        long long ir = (long long)rint((x[i] - xmin) / cellsize);
        long long jr = (long long)rint((y[i] - ymin) / cellsize);
        long long kr = (long long)rint((z[i] - zmin) / cellsize);
        long long c = ir + mx * (jr + (long long)ny * kr);
        next[i] = head[c];
        head[c] = i;
    }
    long long shortq = 0;
    omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : shortq)
    {
        ref_neighbor nb[MAX_NUM_NEAR];
        int npn = 0;
#pragma omp for schedule(dynamic, 1)
        for (long long c = 0; c < ncells; c++) {
            for (long long i = head[c]; i >= 0; i = next[i]) {
                particle pp;
                pp.x = x[i]; pp.y = y[i]; pp.z = z[i];
                for (int j = 0; j < npn; j++) {
                    long long q = nb[j].indx;
                    dptype tx = pp.x - x[q], ty = pp.y - y[q], tz = pp.z - z[q];
                    nb[j].dist2 = tx * tx + ty * ty + tz * tz;
                }
                if (npn) qsort(nb, (size_t)npn, sizeof(ref_neighbor), ref_cmp_nb);
                npn = Find_Near_PreCopied(&pp, k, tree, ptl, nb, npn);
                if (npn != k) shortq++;
                for (int j = 0; j < k; j++) idx[i * k + j] = j < npn ? nb[j].indx : -1;
            }
        }
    }
    free(head);
    free(next);
    free(tree);
    free(ptl);
    return shortq;
}

/* The reference's own findsphdensity (defective Find_Near, nnost.c:375). */
void ref_findsphdensity(const double *x, const double *y, const double *z,
                        const double *m, long long np, int k, long long *idx,
                        float *dens, int nthreads) {
    ref_init(4LL << 30);
    SimpleBasicParticleType *bp = ref_pack(x, y, z, m, NULL, np);
    omp_set_num_threads(nthreads);
    findsphdensity(bp, (int)np, idx, k, dens);
    free(bp);
}

/* The np >= 1e5 neighbour-list builder of lagFindStellarCore as the reference
 * runs it (nnost.c:1361-1529): particles linked to their nearest TSC node
 * cell; shallow cells (fewer than DEEPSIZE members) one per thread with a
 * warm-started list (re-measured, qsort-ed, Find_Near_PreCopied); then each
 * deep cell by all threads, members dealt round-robin, every thread's list
 * seeded with the cell's first k members.  Only cells whose flat index is a
 * multiple of `stride` are run (1: all) -- a bounded sample of the same work
 * for timing.  *nq gets the number of queries run, *t_tree the seconds of the
 * serial tree build.  Returns the number of short lists. */
long long ref_knn_lag(const double *x, const double *y, const double *z, long long np, int k,
                      double xmin, double ymin, double zmin, double cellsize, int nx, int ny,
                      int nz, long long *idx, int nthreads, long long stride, long long *nq,
                      double *t_tree) {
    double t0 = omp_get_wtime();
    TPtlStruct *ptl;
    TStruct *tree = ref_tree(x, y, z, np, &ptl);
    if (t_tree) *t_tree = omp_get_wtime() - t0;
    long long mx = 2 * (nx / 2 + 1);
    long long ncells = mx * (long long)ny * nz;
    long long *head = (long long *)malloc(sizeof(long long) * (size_t)ncells);
    long long *cnt = (long long *)calloc((size_t)ncells, sizeof(long long));
    long long *next = (long long *)malloc(sizeof(long long) * (size_t)np);
    for (long long c = 0; c < ncells; c++) head[c] = -1;
    for (long long i = 0; i < np; i++) {
        long long ir = (long long)rint((x[i] - xmin) / cellsize);
        long long jr = (long long)rint((y[i] - ymin) / cellsize);
        long long kr = (long long)rint((z[i] - zmin) / cellsize);
        long long c = ir + mx * (jr + (long long)ny * kr);
        next[i] = head[c];
        head[c] = i;
        cnt[c]++;
    }
    if (stride < 1) stride = 1;
    long long shortq = 0, nrun = 0;
    omp_set_num_threads(nthreads);
    ref_neighbor *tnb = (ref_neighbor *)malloc(sizeof(ref_neighbor) * (size_t)nthreads * MAX_NUM_NEAR);
#pragma omp parallel reduction(+ : shortq, nrun)
    {
        ref_neighbor *nb = tnb + (size_t)omp_get_thread_num() * MAX_NUM_NEAR;
        int npn = 0;
#pragma omp for schedule(dynamic, 1)
        for (long long c = 0; c < ncells; c += stride) {
            if (cnt[c] == 0 || cnt[c] >= DEEPSIZE) continue;
            for (long long i = head[c]; i >= 0; i = next[i]) {
                particle pp;
                pp.x = x[i]; pp.y = y[i]; pp.z = z[i];
                for (int j = 0; j < npn; j++) {
                    long long q = nb[j].indx;
                    dptype tx = pp.x - x[q], ty = pp.y - y[q], tz = pp.z - z[q];
                    nb[j].dist2 = tx * tx + ty * ty + tz * tz;
                }
                if (npn) qsort(nb, (size_t)npn, sizeof(ref_neighbor), ref_cmp_nb);
                npn = Find_Near_PreCopied(&pp, k, tree, ptl, nb, npn);
                if (npn != k) shortq++;
                nrun++;
                for (int j = 0; j < k; j++) idx[i * k + j] = j < npn ? nb[j].indx : -1;
            }
        }
    }
    for (long long c = 0; c < ncells; c += stride) {
        if (cnt[c] < DEEPSIZE) continue;
        int seed = 0;
        for (long long i = head[c]; i >= 0 && seed < k; i = next[i]) tnb[seed++].indx = i;
        for (int t = 1; t < nthreads; t++)
            for (int j = 0; j < seed; j++) tnb[(size_t)t * MAX_NUM_NEAR + j] = tnb[j];
#pragma omp parallel reduction(+ : shortq, nrun)
        {
            const int tid = omp_get_thread_num(), nth = omp_get_num_threads();
            ref_neighbor *nb = tnb + (size_t)tid * MAX_NUM_NEAR;
            int npn = seed;
            long long icount = 0;
            for (long long i = head[c]; i >= 0; i = next[i], icount++) {
                if (icount % nth != tid) continue;
                particle pp;
                pp.x = x[i]; pp.y = y[i]; pp.z = z[i];
                for (int j = 0; j < npn; j++) {
                    long long q = nb[j].indx;
                    dptype tx = pp.x - x[q], ty = pp.y - y[q], tz = pp.z - z[q];
                    nb[j].dist2 = tx * tx + ty * ty + tz * tz;
                }
                if (npn) qsort(nb, (size_t)npn, sizeof(ref_neighbor), ref_cmp_nb);
                npn = Find_Near_PreCopied(&pp, k, tree, ptl, nb, npn);
                if (npn != k) shortq++;
                nrun++;
                for (int j = 0; j < k; j++) idx[i * k + j] = j < npn ? nb[j].indx : -1;
            }
        }
    }
    if (nq) *nq = nrun;
    free(tnb);
    free(head);
    free(cnt);
    free(next);
    free(tree);
    free(ptl);
    return shortq;
}

/* Cold-start Find_Near_PreCopied for the nq particles qidx[] among all np
 * (tree over all, as ref_knn_precopied).  Returns the number of short lists. */
long long ref_knn_query(const double *x, const double *y, const double *z, long long np,
                        int k, const long long *qidx, long long nq, long long *idx,
                        int nthreads) {
    TPtlStruct *ptl;
    TStruct *tree = ref_tree(x, y, z, np, &ptl);
    long long shortq = 0;
    omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : shortq)
    {
        ref_neighbor nb[MAX_NUM_NEAR];
#pragma omp for schedule(dynamic, 16)
        for (long long t = 0; t < nq; t++) {
            const long long i = qidx[t];
            particle pp;
            pp.x = x[i]; pp.y = y[i]; pp.z = z[i];
            int n = Find_Near_PreCopied(&pp, k, tree, ptl, nb, 0);
            if (n != k) shortq++;
            for (int j = 0; j < k; j++) idx[t * k + j] = j < n ? nb[j].indx : -1;
        }
    }
    free(tree);
    free(ptl);
    return shortq;
}

/* The reference's finddenpeak (subhaloden.mod6.c:416-492) on densities and
 * neighbour lists given by the caller: cores (peak particle, its position as
 * the record's floats, its density) copied out, at most maxout of them.
 * Returns the number of cores. */
int finddenpeak(float *, int, long long *, int, Coretype **, int, SimpleBasicParticleType *);

long long ref_finddenpeak(const float *den, int k, const long long *nbr, long long np,
                          const double *x, const double *y, const double *z, const int *type,
                          int jflag, int *peak, float *cx, float *cy, float *cz, float *cden,
                          long long maxout) {
    ref_init(4LL << 30);
    SimpleBasicParticleType *bp = ref_pack(x, y, z, NULL, type, np);
    Coretype *core = (Coretype *)Malloc(sizeof(Coretype) * MAXNUMCORE, PPTR(core));
    int n = finddenpeak((float *)den, k, (long long *)nbr, (int)np, &core, jflag, bp);
    for (long long i = 0; i < n && i < maxout; i++) {
        peak[i] = core[i].peak;
        cx[i] = core[i].cx;
        cy[i] = core[i].cy;
        cz[i] = core[i].cz;
        cden[i] = core[i].density;
    }
    Free(core);
    free(bp);
    return n;
}

/* MergingPeak (subhaloden.mod6.c:341-415, friends-of-friends of ost.c) on
 * cores given by their peak particle and density (iflag 0: ordered by
 * density); returns the new core count.  order_peak[0 .. numcore) receives
 * the peaks of the whole core array after the call (the kept cores first,
 * then what the reference leaves behind them). */
int MergingPeak(SimpleBasicParticleType *, int, Coretype *, int, int);

int ref_merging_peak(const double *x, const double *y, const double *z, long long np,
                     const int *peak, const float *density, int numcore, int *order_peak) {
    ref_init(4LL << 30);
    SimpleBasicParticleType *bp = ref_pack(x, y, z, NULL, NULL, np);
    Coretype *core = (Coretype *)calloc((size_t)numcore, sizeof(Coretype));
    for (int i = 0; i < numcore; i++) {
        core[i].peak = peak[i];
        core[i].density = density[i];
    }
    int n = MergingPeak(bp, (int)np, core, numcore, 0);
    for (int i = 0; i < numcore; i++) order_peak[i] = core[i].peak;
    free(core);
    free(bp);
    return n;
}

/* The watershed of subhalo_den (subhaloden.mod6.c:2978-3024) with the
 * reference's own Dump2WorkingParticle, FindCoreDensity,
 * GetShellParticlesPeaks and SaveRemainingParticles2LastShell.  They keep
 * their state in file-scope globals of subhaloden.mod6.c: the working
 * particles `wp`, the shells `shell[]` and their count `nshell`.  The two
 * record types are file-local there (:237-242, :306-311) and restated here
 * with the same members, since this driver fills and reads them.
 * Cores are given by their peak particle (core order as given); den is
 * wp[].den.  FindCoreDensity runs on `nthreads` OpenMP threads.  With
 * with_shells, the shell thresholds thr[0 .. nthr) are applied in order and
 * the shells copied out as CSR (particles, cores); returns nshell, -1
 * when more than maxshell, -3 when the shell cores exceed maxcores.
 * Without, returns 0. */
typedef struct {
    float den;
    int haloid;
    unsigned char flag;
    unsigned char flag2[MAXTHREADS];
} ref_working_particle;
typedef struct {
    int nlinkedcore;
    int *linkedcore;
    int nlinkedparticle;
    int *linkedparticle;
} ref_shell_type;
extern ref_working_particle *wp;
extern ref_shell_type shell[];
extern int nshell;
void Dump2WorkingParticle(float *, int, Coretype *, int);
int FindCoreDensity(SimpleBasicParticleType *, int, long long *, int, Coretype *, int);
void GetShellParticlesPeaks(int, long long *, int, Coretype *, int, float);
void SaveRemainingParticles2LastShell(int, Coretype *, int);

long long ref_core_shells(const int *type, const double *mass, long long np, const float *den,
                          const long long *nbr, int k, const int *peak, int numcore,
                          int nthreads, float *coredensity, int *nummem, int *numstar,
                          float *starmass, int *haloid, unsigned char *flag, int with_shells,
                          const float *thr, int nthr, long long *part_ptr, int *part,
                          long long *core_ptr, int *cores, long long maxshell,
                          long long maxcores) {
    ref_init(4LL << 30);
    SimpleBasicParticleType *bp =
        (SimpleBasicParticleType *)calloc((size_t)np, sizeof(SimpleBasicParticleType));
    for (long long i = 0; i < np; i++) {
        bp[i].type = (unsigned int)type[i];
        bp[i].mass = mass[i];
        bp[i].indx = (idtype)i;
    }
    Coretype *core = (Coretype *)calloc((size_t)(numcore > 0 ? numcore : 1), sizeof(Coretype));
    for (int c = 0; c < numcore; c++) core[c].peak = peak[c];
    wp = (ref_working_particle *)Malloc(sizeof(ref_working_particle) * np, PPTR(wp));
    Dump2WorkingParticle((float *)den, (int)np, core, numcore);
    omp_set_num_threads(nthreads);
    FindCoreDensity(bp, (int)np, (long long *)nbr, k, core, numcore);
    for (int c = 0; c < numcore; c++) {
        coredensity[c] = core[c].coredensity;
        nummem[c] = core[c].nummem;
        numstar[c] = core[c].numstar;
        starmass[c] = core[c].starmass;
    }
    long long ret = 0;
    if (with_shells) {
        nshell = 0;
        for (int t = 0; t < nthr; t++)
            GetShellParticlesPeaks((int)np, (long long *)nbr, k, core, numcore, thr[t]);
        SaveRemainingParticles2LastShell((int)np, core, numcore);
        ret = nshell <= maxshell ? nshell : -1;
        long long np_out = 0, nc_out = 0, nc_all = 0;
        for (int s = 0; s < nshell; s++) nc_all += shell[s].nlinkedcore;
        if (ret >= 0 && nc_all > maxcores) ret = -3;
        if (ret >= 0) {
            part_ptr[0] = core_ptr[0] = 0;
            for (int s = 0; s < nshell; s++) {
                for (int j = 0; j < shell[s].nlinkedparticle; j++)
                    part[np_out++] = shell[s].linkedparticle[j];
                for (int j = 0; j < shell[s].nlinkedcore; j++)
                    cores[nc_out++] = shell[s].linkedcore[j];
                part_ptr[s + 1] = np_out;
                core_ptr[s + 1] = nc_out;
            }
        }
        for (int s = nshell - 1; s >= 0; s--) {
            Free(shell[s].linkedcore);
            Free(shell[s].linkedparticle);
            shell[s].linkedcore = NULL;
            shell[s].linkedparticle = NULL;
            shell[s].nlinkedcore = shell[s].nlinkedparticle = 0;
        }
        nshell = 0;
    }
    for (long long i = 0; i < np; i++) {
        haloid[i] = wp[i].haloid;
        flag[i] = wp[i].flag;
    }
    Free(wp);
    free(core);
    free(bp);
    return ret;
}
