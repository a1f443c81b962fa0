/*
 * pgalf_density.h -- C-ABI of the MI355X (gfx950) density stage of
 * NewGalFinder (kjhan0606/GalaxyFinder).
 *
 * Part 1 exports the reference's own symbol names with the reference's
 * argument lists, so the galaxy finder links against this library in place
 * of tsc_omp2.o, gsmooth.o and the kNN density routines of nnost.o.  The
 * calls are synchronous: inputs are host arrays, outputs are written to the
 * caller's host buffers before return.  They never exit(); failures are
 * reported through pgalf_last_error().
 *
 * Part 2 is the device-resident API (float32 SoA device pointers, a HIP
 * stream) used by the Python/PyTorch wrapper and the benchmarks.
 */
#ifndef PGALF_DENSITY_H
#define PGALF_DENSITY_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as the reference's SimpleBasicParticleType (tree.h:155-162,
 * 88 bytes under the reference Makefile flags): type@0, mass@8, x,y,z@16,24,
 * 32, vx,vy,vz@40,48,56, link02@64, indx@72, bp@80. */
typedef struct pgalf_particle {
    unsigned int type;
    double mass, x, y, z;
    double vx, vy, vz;
    double link02;
    int indx;
    struct pgalf_particle *bp;
} pgalf_particle_t;

/* Particle types (tree.h:47-48). */
#define PGALF_TYPE_STAR 1
#define PGALF_TYPE_DM 2
#define PGALF_TYPE_GAS 3
#define PGALF_TYPE_SINK 4
#define PGALF_TYPE_ALL 5

/* Status codes. */
#define PGALF_OK 0
#define PGALF_ERR_ARG 1          /* bad argument (sizes, k > 32, k > np ...) */
#define PGALF_ERR_HIP 2          /* HIP runtime error (no device, launch ...) */
#define PGALF_ERR_OUT_OF_GRID 3  /* a TSC stencil left the grid; those
                                    particles were skipped (the reference
                                    reads/writes out of bounds there) */
#define PGALF_ERR_NOMEM 4        /* device allocation failed */
#define PGALF_ERR_SHORT_LIST 5   /* fewer than k candidates (reference: exit(99)) */
#define PGALF_ERR_CORE_OVERFLOW 6 /* more cores than the caller's array holds
                                     (reference: exit(999) or Realloc) */

/* Same layout as the reference's Coretype (tree.h:204-213, 56 bytes).  The
 * peak entry points fill peak, cx, cy, cz and density (as the reference does)
 * and zero the other fields. */
typedef struct pgalf_core {
    int peak;
    int nummem, numstar;
    float starmass;
    float coredensity;
    float cx, cy, cz;
    float cvx, cvy, cvz;
    float Rtidal, density;
    unsigned char flag;
} pgalf_core_t;

/* Reference constants of the peak stage (params.h:19,115). */
#define PGALF_PEAKTHRESHOLD 2000.0f
#define PGALF_NCELLBUFF 10

/* Status of the last call on this host thread, and its message. */
int pgalf_last_error(void);
const char *pgalf_last_error_string(void);

/* ===================== Part 1: reference entry points ===================== */

/* TSC mass deposit onto the padded float grid den[mx*ny*nz], mx=2*(nx/2+1);
 * particles with type != ptype are skipped unless ptype == PGALF_TYPE_ALL.
 * Replaces assign_density_TSC (NewGalFinder/tsc_omp2.c:166). */
void assign_density_TSC(pgalf_particle_t *bp, int np, float *den, int nx, int ny,
                        int nz, double Xmin, double Ymin, double Zmin,
                        double cellsize, int ptype);

/* In-place FFT Gaussian smoothing of the padded grid, kernel exp(-r^2/RG^2).
 * Replaces gaussian_Smoothing (NewGalFinder/gsmooth.c:14). */
void gaussian_Smoothing(float *denGrid, int nx, int ny, int nz, double cellsize,
                        float RG);

/* TSC read-back of the grid at every particle (all types) into densph[np].
 * Replaces findDen (NewGalFinder/tsc_omp2.c:40). */
void findDen(pgalf_particle_t *bp, int np, float *densph, double xmin, double ymin,
             double zmin, double cellsize, float *den, int nx, int ny, int nz);

/* Exact Numnear-NN lists (self included, ascending distance, ties by index)
 * into nearindex[np*Numnear] and the unweighted W4 kNN density densph[np].
 * Replaces findsphdensity (NewGalFinder/nnost.c:560). */
void findsphdensity(pgalf_particle_t *bp, int np, long long *nearindex, int Numnear,
                    float *densph);

/* Mass-weighted W4 density over the 10 nearest stars (NUMNEARDEN,
 * params.h:73) into densph[np], and all-particle Numnear-NN lists.
 * Replaces starfindsphdensity (NewGalFinder/nnost.c:632). */
void starfindsphdensity(pgalf_particle_t *bp, int np, long long *nearindex,
                        int Numnear, float *densph);

/* The all-particle Numnear-NN lists that lagFindStellarCore builds inline
 * (NewGalFinder/nnost.c:1298-1536, its Find_Near / warm-started
 * Find_Near_PreCopied loops).  Returns a status code. */
int pgalf_near_lists(pgalf_particle_t *bp, int np, int Numnear, long long *nearindex);

/* Whole density stage of lagFindStellarCore up to the neighbour lists
 * (nnost.c:1030-1096 and 1298-1536) in one device residency: grid box and
 * size as nnost.c:1040-1058 (TSC_CELL_SIZE 0.004, 10-cell buffer), deposit
 * of `ptype` particles, smoothing with RG = 0.008, read-back into densph[np]
 * and the Numnear-NN lists.  The smoothed grid is returned in *grid_out
 * (malloc'ed, mx*ny*nz floats; free() it) with its geometry, if non-null.
 * Returns a status code. */
int pgalf_lag_density_stage(pgalf_particle_t *bp, int np, int Numnear, int ptype,
                            float *densph, long long *nearindex, float **grid_out,
                            int *nxyz_out, double *xyzmin_out);

/* Grid cores of lagFindStellarCore (nnost.c:1115-1271, the block between the
 * read-back and MergingPeak): particles (stars for ptype PGALF_TYPE_STAR, all
 * for PGALF_TYPE_ALL) on their nearest node of the smoothed grid denGrid;
 * every node at least nbuff/2 cells inside the grid whose value exceeds
 * `threshold` and its 26 neighbours and which holds particles gives a core:
 * its particle of largest densph (the largest index among equal ones),
 * skipped when particle 0 is among the node's particles (the reference's
 * `if(ibp > 0)`).  Cores in (k, j, i) node order into core[0 .. *numcore);
 * PGALF_ERR_CORE_OVERFLOW when more than maxcore (then *numcore is the
 * count).  The reference uses threshold = PGALF_PEAKTHRESHOLD and
 * nbuff = PGALF_NCELLBUFF.  Returns a status code. */
int pgalf_grid_cores(pgalf_particle_t *bp, int np, const float *densph, const float *denGrid,
                     int nx, int ny, int nz, double xmin, double ymin, double zmin,
                     double cellsize, int ptype, float threshold, int nbuff,
                     pgalf_core_t *core, int maxcore, int *numcore);

/* finddenpeak (subhaloden.mod6.c:416-492) with a caller-sized core array in
 * place of the reference's Realloc'ed one: particle i is a core when den[i] >
 * PGALF_PEAKTHRESHOLD, it is a star if jflag == 1, and no particle of its
 * list neighbor[i*numneigh ..] is strictly denser.  Cores in index order into
 * core[0 .. *numcore); PGALF_ERR_CORE_OVERFLOW when more than maxcore.
 * Returns a status code. */
int pgalf_finddenpeak(float *den, int numneigh, long long *neighbor, int np, int jflag,
                      pgalf_particle_t *bp, pgalf_core_t *core, int maxcore, int *numcore);

/* MergingPeak (subhaloden.mod6.c:341-415, same arguments and return value):
 * the cores are ordered by density (iflag 0) or by numstar (iflag 1),
 * largest first, equal keys in their input order; of each
 * friends-of-friends group of their peak particles (link length
 * PGALF_MERGINGPEAKLENGTH, ost.c:306-356) the first core is kept.  core[0 ..
 * return) holds the kept cores in that order, core[return .. numcore) the
 * rest of the ordered array, as the reference leaves it.  iflag 1 does not
 * touch the reference's static peak flags (subhaloden.mod6.c:199-212, which
 * its only caller with iflag 1 maintains, :1465).  numcore <= 0 returns 0;
 * on error -1 (pgalf_last_error). */
#define PGALF_MERGINGPEAKLENGTH 4.0e-3f /* params.h:49 */
int pgalf_merging_peak(pgalf_particle_t *bp, int np, pgalf_core_t *core, int numcore,
                       int iflag);

/* Watershed of subhalo_den.  The reference keeps the working particles in a
 * global array `wp` (subhaloden.mod6.c:237-243); here its fields are three
 * caller arrays of np entries: den (wp[].den as Dump2WorkingParticle sets
 * it, :493-500), haloid (wp[].haloid) and flag (wp[].flag, bits below).
 * Flag bits (subhaloden.mod6.c:180-186) and the non-member id (defs.h:29): */
#define PGALF_WP_PEAK 0x01
#define PGALF_WP_VISITED 0x02
#define PGALF_WP_CORE 0x04
#define PGALF_WP_SHELL 0x08
#define PGALF_WP_BOUND 0x10
#define PGALF_WP_REMAINING 0x20
#define PGALF_NOT_HALO_MEMBER (-9999)
#define PGALF_MAX_SHELL_DIVIDE 15 /* the most shell thresholds subhalo_den uses */

/* FindCoreDensity (subhaloden.mod6.c:1167-1475; same arguments, same return
 * value numcore) with wp as den/haloid/flag.  Reads core[c].peak and
 * bp[].type/mass; writes core[c].coredensity, nummem, numstar, starmass (the
 * FP32 sum of the member stars' masses in sweep order), flag (PEAK, CORE,
 * BOUND) and haloid (the core of each member, PGALF_NOT_HALO_MEMBER
 * elsewhere).  The reference spreads the cores over OpenMP threads; this is
 * the one-thread result: a particle among the members of several cores has
 * the last core's id.  NumNeighbor <= 32, np < 2^26.  On error -1
 * (pgalf_last_error). */
int pgalf_find_core_density(pgalf_particle_t *bp, int np, long long *neighbor, int NumNeighbor,
                            pgalf_core_t *core, int numcore, const float *den, int *haloid,
                            unsigned char *flag);

/* Shells as CSR arrays owned by the caller: shell s holds particles
 * part[part_ptr[s] .. part_ptr[s+1]) and cores cores[core_ptr[s] ..
 * core_ptr[s+1]) (the reference's SHELL2P(s) / SHELL2C(s),
 * subhaloden.mod6.c:285-295).  part holds np entries, part_ptr and core_ptr
 * maxshell + 1; a core listed by several shells of one threshold appears in
 * each, so cores may need more than numcore * nthr entries. */
typedef struct pgalf_shells {
    long long *part_ptr;
    int *part;
    long long *core_ptr;
    int *cores;
    long long maxshell, maxcores;
    long long nshell; /* out */
} pgalf_shells_t;

/* The shell thresholds of subhalo_den (subhaloden.mod6.c:2995-3024):
 * nshelldivide values (10, or up to 15 for np > 1000), highest first, into
 * thr[0 .. return).  On error -1. */
int pgalf_shell_thresholds(const float *den, int np, const pgalf_core_t *core, int numcore,
                           float *thr);

/* GetShellParticlesPeaks (subhaloden.mod6.c:612-678) for each of thr[0 ..
 * nthr) in order, then SaveRemainingParticles2LastShell (:572-611), on the
 * state FindCoreDensity left (flag, haloid updated in place; core[c].peak
 * and coredensity read).  The shells go to *shells (nshell set).  Returns a
 * status code: PGALF_ERR_CORE_OVERFLOW when the shells or their core entries
 * exceed maxshell / maxcores, PGALF_ERR_ARG when two cores share a peak
 * (the reference never ends there). */
int pgalf_shell_particles_peaks(int np, long long *neighbor, int NumNeighbor, pgalf_core_t *core,
                                int numcore, const float *den, int *haloid, unsigned char *flag,
                                const float *thr, int nthr, pgalf_shells_t *shells);

/* ===================== Part 2: device-resident API ======================== */

typedef struct pgalf_grid {
    int nx, ny, nz;
    double xmin, ymin, zmin, cellsize;
} pgalf_grid_t;

/* All pointers are device pointers; `stream` is a hipStream_t (NULL for the
 * default stream).  Calls are stream-ordered; their status is that of the
 * launch.  Calls share one scratch arena per device: a call on a stream other
 * than the previous call's first makes its stream wait for the previous
 * call's kernels (an event), so calls on different streams are serialised on
 * the device rather than overlapped.  Kernel-detected conditions (out-of-grid particles, short lists)
 * are collected with pgalf_dev_status().  Grids that a call writes (`den` of
 * pgalf_dev_tsc, pgalf_dev_smooth, pgalf_dev_density_stage and the batch
 * runs) must be 8-byte aligned: they are written as float pairs
 * (PGALF_ERR_ARG otherwise). */
int pgalf_dev_tsc(const float *x, const float *y, const float *z, const float *m,
                  const int *type, long long np, int ptype, const pgalf_grid_t *g,
                  float *den, void *stream);
int pgalf_dev_smooth(float *den, const pgalf_grid_t *g, float RG, void *stream);
int pgalf_dev_findden(const float *x, const float *y, const float *z, long long np,
                      const pgalf_grid_t *g, const float *den, float *densph,
                      void *stream);
/* k <= 32 nearest of each particle among the particles whose type equals
 * cand_ptype (PGALF_TYPE_ALL or type == NULL: all).  rho_mode 0: none,
 * 1: unweighted W4 density, 2: mass-weighted.  nbr/rho may be NULL. */
int pgalf_dev_knn(const float *x, const float *y, const float *z, const float *m,
                  const int *type, long long np, int cand_ptype, int k,
                  long long *nbr, float *rho, int rho_mode, void *stream);
/* pgalf_dev_knn without any allocation (no counterpart in the reference; for
 * callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * pgalf_dev_knn; its scratch is a static 64 MiB device array, one per GPU,
 * held until the process ends, so np is limited to PGALF_KNN_NOALLOC_MAX
 * (PGALF_ERR_ARG beyond it).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other host threads' streams)
 * race and must be ordered by the caller. */
#define PGALF_KNN_NOALLOC_MAX 524288LL
int pgalf_dev_knn_noalloc(const float *x, const float *y, const float *z, const float *m,
                          const int *type, long long np, int cand_ptype, int k,
                          long long *nbr, float *rho, int rho_mode, void *stream);
long long pgalf_dev_knn_noalloc_max(void);
/* deposit -> smooth -> read-back -> kNN lists + W4 density, one call.  With
 * a kNN to run and `stream` not under graph capture, the deposit, smoothing
 * and read-back run on a stream of the library's own (default priority)
 * beside the kNN on `stream`, ordered by events: they start after the work
 * queued on `stream` before the call, and `stream` waits for them before
 * anything queued after it.  Under capture the stages run one after the
 * other on `stream`.  The results are the same either way. */
int pgalf_dev_density_stage(const float *x, const float *y, const float *z,
                            const float *m, long long np, const pgalf_grid_t *g,
                            float RG, int k, float *den, float *rho_tsc,
                            long long *nbr, float *rho_knn, void *stream);
/* Adjoint of the stage (no counterpart in the reference; added for
 * differentiable density estimates).  With the forward's smoothed grid `den`
 * (from pgalf_dev_tsc + pgalf_dev_smooth with the same grid and RG) and
 * neighbour lists `nbr` (pgalf_dev_knn, indices into the same arrays), and
 * the cotangents g_tsc = dL/drho_tsc (read-back density; NULL: none) and
 * g_knn = dL/drho_knn (W4 kNN density of rho_mode 1 unweighted / 2
 * mass-weighted; NULL: none), writes dL/dx, dL/dy, dL/dz, dL/dm (float32,
 * overwritten; any may be NULL).  Derivatives of the FP64 mathematical
 * forward (TSC weights, W4 with h = r_k/2) with the lists held fixed;
 * accumulated in FP64. */
int pgalf_dev_density_stage_backward(const float *x, const float *y, const float *z,
                                     const float *m, long long np, const pgalf_grid_t *g,
                                     float RG, int k, const float *den,
                                     const long long *nbr, int rho_mode, const float *g_tsc,
                                     const float *g_knn, float *gx, float *gy, float *gz,
                                     float *gm, void *stream);
/* Peak climbing on a smoothed grid `den` (no counterpart in the reference;
 * uses the density gradients of the adjoint machinery): niter normalised
 * gradient-ascent steps of `step` cells on the TSC read-back field of den
 * (positions kept within [1, n-2] cells), climbed positions to ox, oy, oz,
 * then an exact integer steepest ascent over the 26-neighbourhood of den
 * from the nearest node to a local maximum: its unpadded flat index
 * i + nx*(j + ny*k) is the particle's peak (membership). */
int pgalf_dev_peak_climb(const float *x, const float *y, const float *z, long long np,
                         const pgalf_grid_t *g, const float *den, int niter, float step,
                         float *ox, float *oy, float *oz, int *peak, void *stream);
/* Device forms of pgalf_grid_cores and pgalf_finddenpeak: core particle
 * indices to core[0 .. min(*ncore, maxcore)), the count *ncore left in device
 * memory (stream-ordered, no synchronisation). */
int pgalf_dev_grid_cores(const float *x, const float *y, const float *z, const int *type,
                         long long np, int ptype, const float *densph, const pgalf_grid_t *g,
                         const float *den, float threshold, int nbuff, int *core,
                         long long maxcore, int *ncore, void *stream);
int pgalf_dev_finddenpeak(const float *den, int k, const long long *nbr, long long np,
                          const int *type, int jflag, float threshold, int *core,
                          long long maxcore, int *ncore, void *stream);
/* Device form of pgalf_merging_peak: core i at particle peak[i], ranked by
 * key[i] (largest first, ties in input order); keep[0 .. *nkeep) receives
 * the input indices of the kept cores in rank order, order (nullable) the
 * input index of every rank; *nkeep is left in device memory.  link: the
 * reference uses PGALF_MERGINGPEAKLENGTH.  numcore < 2^30.  A peak outside
 * [0, np) links to nothing and is reported by pgalf_dev_status. */
int pgalf_dev_merging_peak(const float *x, const float *y, const float *z, long long np,
                           const int *peak, const float *key, int numcore, float link,
                           int *keep, int *order, int *nkeep, void *stream);
/* Device forms of the watershed (den, nbr, peak, type, mass, outputs and
 * the shell arrays in device memory; thr in host memory).  They read counts
 * back between stages, so they synchronise `stream` and cannot be captured
 * in a graph.  mass: the particles' FP64 masses (starmass sums them). */
int pgalf_dev_find_core_density(const float *den, long long np, const long long *nbr, int k,
                                const int *type, const double *mass, const int *peak,
                                int numcore, float *coredensity, int *nummem, int *numstar,
                                float *starmass, int *haloid, unsigned char *flag, void *stream);
int pgalf_dev_shell_thresholds(const float *den, long long np, const float *coredensity,
                               int numcore, float *thr, int *nthr, void *stream);
int pgalf_dev_shell_particles_peaks(const float *den, long long np, const long long *nbr, int k,
                                    const int *peak, const float *coredensity, int numcore,
                                    const float *thr, int nthr, unsigned char *flag,
                                    int *haloid, pgalf_shells_t *shells, void *stream);
/* Synchronise `stream` and return the status of the kernels run since the
 * previous call (PGALF_ERR_OUT_OF_GRID / PGALF_ERR_SHORT_LIST / PGALF_ERR_ARG
 * for an index outside the particle range, or PGALF_OK). */
int pgalf_dev_status(void *stream);

/* ===================== Part 3: batched catalogue API ====================== */

/* A catalogue of halos run through the density stage together, in place of
 * the one-halo-per-call loop of the reference's MPI workers
 * (NewGalFinder/gfind.c:394-549 dispatching subhalo_den per halo).
 * The particles of halo h are [off[h], off[h+1]) of the concatenated float32
 * SoA device arrays (off[0] == 0, each halo >= k particles).  Halo h is
 * deposited onto its own padded grid grids[h] (at float offset
 * pgalf_batch_grid_offset(b, h) of one batch grid buffer of
 * pgalf_batch_grid_offset(b, nhalo) floats), smoothed with RG[h] (box
 * units, the reference's gaussian_Smoothing argument), read back, and gets
 * its exact k-NN lists (indices local to the halo, as nearindex is in
 * lagFindStellarCore) and unweighted W4 densities.  The plan (descriptors,
 * FFT plans, tree layout, wave table) is built once per catalogue shape.
 * Returns NULL on failure (see pgalf_last_error()). */
typedef struct pgalf_batch pgalf_batch_t;
pgalf_batch_t *pgalf_batch_create(int nhalo, const long long *off, const pgalf_grid_t *grids,
                                  const float *RG, int k);
long long pgalf_batch_grid_offset(const pgalf_batch_t *b, int h);
/* One stage pass over the catalogue, stream-ordered on `stream`.  rho_tsc,
 * nbr and rho_knn may be NULL.  use_graph != 0: the launch sequence is
 * captured into a HIP graph on first use (per set of pointers) and replayed
 * afterwards. */
int pgalf_batch_run(pgalf_batch_t *b, const float *x, const float *y, const float *z,
                    const float *m, float *den, float *rho_tsc, long long *nbr,
                    float *rho_knn, int use_graph, void *stream);
/* One differentiable pass over the catalogue with peak climbing (no
 * counterpart in the reference), stream-ordered and, with use_graph != 0,
 * replayed from one HIP graph:
 *  - forward as pgalf_batch_run, with the W4 kNN density of rho_mode
 *    1 (unweighted) or 2 (mass-weighted);
 *  - the cotangents of L = sum w_tsc log rho_tsc + sum w_knn log rho_knn
 *    (w_tsc / w_knn may be NULL: that term is absent);
 *  - the adjoint (as pgalf_dev_density_stage_backward) into gx, gy, gz, gm;
 *  - peak climbing on each halo's smoothed grid (as pgalf_dev_peak_climb):
 *    climbed positions in cx, cy, cz and each particle's peak cell (its
 *    membership) in peak.
 * All other pointers are required. */
int pgalf_batch_run_grad(pgalf_batch_t *b, const float *x, const float *y, const float *z,
                         const float *m, const float *w_tsc, const float *w_knn, int rho_mode,
                         float *den, float *rho_tsc, long long *nearindex, float *rho_knn,
                         float *gx, float *gy, float *gz, float *gm, int climb_iters,
                         float climb_step, float *cx, float *cy, float *cz, int *peak,
                         int use_graph, void *stream);
/* Synchronise `stream` and return the status of the batch's kernels since
 * the previous call (PGALF_ERR_OUT_OF_GRID or PGALF_OK). */
int pgalf_batch_status(pgalf_batch_t *b, void *stream);
/* Device scratch of the batch (no counterpart in the reference; for
 * capacity planning): reserved == 0, the largest scratch any of its runs has
 * held at once; reserved != 0, the bytes its arena holds now.  -1 for a
 * null batch. */
long long pgalf_batch_workspace_bytes(const pgalf_batch_t *b, int reserved);

/* The watershed of subhalo_den over a whole catalogue: FindCoreDensity, the
 * shell thresholds and the shells (GetShellParticlesPeaks for each,
 * SaveRemainingParticles2LastShell) of every halo, each exactly as
 * pgalf_dev_find_core_density / pgalf_dev_shell_particles_peaks give them on
 * the halo alone.  The reference runs these per halo in each MPI worker's
 * subhalo_den (subhaloden.mod6.c:2978-3024); there is no one reference
 * function for the catalogue.  Halo h: particles [off[h], off[h+1]) of the
 * concatenated device arrays den, type, mass (FP64) and of the per-particle
 * outputs haloid (a core index local to the halo, or PGALF_NOT_HALO_MEMBER),
 * flag; its k-entry lists in nbr hold indices local to the halo (as
 * pgalf_batch_run writes them); its cores [core_off[h], core_off[h+1]) of
 * peak (local particle indices) and of the per-core outputs.  off and
 * core_off are host arrays of nhalo + 1 entries.
 * Shells of halo h: shell_base[h], core_base[h] and core_cap[h] from
 * pgalf_catalog_shell_layout; its nshell[h] shells are part_ptr[shell_base[h]
 * .. + nshell[h] + 1) (offsets into its particles part[off[h] ..), local
 * indices) and core_ptr (the same range; offsets into cores[core_base[h]
 * ..), local core indices).  Halos of at most small_max particles (0:
 * 131072) run one per workgroup in one launch per phase, on a stream of the
 * library's own; larger ones (largest first) through the one-halo path, on
 * `stream` and on three more streams of the library's, each driven by a
 * host thread with its own scratch, all at the same time.  The results do
 * not depend on small_max or on that split.  nshell and halo_status are
 * host arrays of nhalo entries;
 * halo_status[h] != 0 marks a halo whose results are not valid (2: a
 * core-density search did not converge; 13: its core entries outgrew
 * core_cap[h]); the call then returns PGALF_ERR_CORE_OVERFLOW or
 * PGALF_ERR_ARG.  Synchronises `stream`.
 * pgalf_catalog_shell_layout fills the three host arrays (nhalo entries)
 * and returns the number of part_ptr / core_ptr entries; *total_cores gets
 * the size of cores.  Halo h has cores x nshelldivide + 2 pointer entries
 * and core_cap[h] = 16 x cores x (nshelldivide + 1) + 16 core entries (a
 * core is listed by every shell sweep that reaches it). */
long long pgalf_catalog_shell_layout(int nhalo, const long long *off, const long long *core_off,
                                     long long *shell_base, long long *core_base,
                                     long long *core_cap, long long *total_cores);
int pgalf_dev_catalog_watershed(int nhalo, const long long *off, const long long *core_off,
                                const float *den, const long long *nbr, int k, const int *type,
                                const double *mass, const int *peak, float *coredensity,
                                int *nummem, int *numstar, float *starmass, int *haloid,
                                unsigned char *flag, long long *part_ptr, int *part,
                                long long *core_ptr, int *cores, long long small_max,
                                long long *nshell, int *halo_status, void *stream);
void pgalf_batch_destroy(pgalf_batch_t *b);

/* Launch schedule (no counterpart in the reference): the knobs an autotuner
 * sets, per bin of workloads up to nmax particles (bins by ascending nmax;
 * a single-halo call picks the bin of its particle count, a batch the bin of
 * its total when it is created; beyond the last bin, the defaults).  Only
 * smooth_direct_max changes result bits, and its methods agree within the
 * smoothing tolerance; the kNN lists and the deposit are identical under
 * every schedule. */
typedef struct pgalf_schedule {
    long long nmax;        /* bin: workloads of at most nmax particles */
    int knn_packed;        /* 1 (default): packed-key kNN query; 0: exact query only */
    int knn_xcd_chunk;     /* consecutive kNN query waves per XCD run, 0..1024 (64) */
    int tsc_unit_entries;  /* bin entries per deposit work unit, 256..65536 (4096) */
} pgalf_schedule_t;
/* Replace the table (nbins 0: the defaults) and the global smoothing
 * threshold: lines up to smooth_direct_max long (0..64, default 56; rows
 * also <= 56) are smoothed by direct convolution, longer ones by FFT. */
int pgalf_set_schedule(const pgalf_schedule_t *bins, int nbins, int smooth_direct_max);
/* The schedule a workload of n particles gets, and the smoothing threshold. */
int pgalf_get_schedule(long long n, pgalf_schedule_t *out, int *smooth_direct_max);

/* Phase timing with HIP events on the launch stream (off by default). */
void pgalf_timing_enable(int on);
void pgalf_timing_reset(void);
/* Sum of elapsed ms and number of intervals recorded for `phase`
 * ("tsc", "smooth", "findden", "knn", "knn_query", ...). */
int pgalf_timing_get(const char *phase, double *total_ms, int *count);

/* Library version string. */
const char *pgalf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PGALF_DENSITY_H */
