/*
 * pgalf_oracle.c -- CPU restatement of NewGalFinder's per-halo density stage.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library in galaxyfinder_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product path never does.
 *
 * Every function restates the reference algorithm with the reference's own
 * arithmetic (operation order, float/double conversions, round-half-even
 * node choice), citing the reference file:line it follows.  No reference
 * source text is copied; the formulas are re-derived and re-written.
 *
 * Pins (see tests/test_oracle_*.py):
 *   orc_tsc_deposit_f32   -- bit-exact vs the reference build (oracle/_ref)
 *                            on committed golden grids, any thread count;
 *                            partition of unity; single-particle weights.
 *   orc_tsc_deposit_f64   -- same contributions, FP64 sums (error reference).
 *   orc_findden           -- bit-exact vs the reference build; constant and
 *                            linear fields reproduced; adjoint identity.
 *   orc_gsmooth           -- reference gsmooth.c is unbuildable (no FFTW in
 *                            the image): pinned by plane-wave eigenvalues,
 *                            DC preservation and the delta response written
 *                            as a closed-form finite cosine sum.
 *   orc_knn_*             -- brute force vs k-d tree; vs the reference
 *                            Find_Near_PreCopied lists in golden fixtures.
 *   orc_W4, orc_sph_*     -- vs the reference W4 (golden table) and closed
 *                            forms W(0,h)=1/(pi h^3), W(2h)=0.
 *   orc_finddenpeak       -- bit-exact vs the reference build's finddenpeak
 *                            (subhaloden.mod6.c compiled in place) on golden
 *                            densities and lists.
 *   orc_merging_peak      -- vs the reference build's MergingPeak (with
 *                            ost.c's friends-of-friends tree, compiled in
 *                            place) on golden core sets; brute-force
 *                            component invariants.
 *   orc_grid_cores        -- the reference block is inline in
 *                            lagFindStellarCore (unbuildable: it calls the
 *                            FFTW smoothing): pinned by an independent
 *                            vectorised brute force and dedicated cases
 *                            (plateaus, ties, the particle-0 skip).
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp, no FMA contraction).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TYPE_ALL 5

/* Reference volume factor: (0.001/TSC_CELL_SIZE)^3 with TSC_CELL_SIZE=0.004,
 * a compile-time constant independent of the cellsize argument
 * (tsc_omp2.c:188, params.h:92). */
static double orc_volfact(void) {
    double r = 0.001 / 0.004;
    return r * r * r;
}

/* ------------------------------------------------------------------------ */
/* TSC weights, deposit flavour (tsc_omp2.c:303-353).                       */
/* The x weights carry the particle mass; y/z weights do not.               */
/* ------------------------------------------------------------------------ */
typedef struct {
    long long near[3];
    long long sign[3];
    float w[3][3]; /* w[axis][0]=centre, [1]=near+sign, [2]=near-sign */
} orc_tscw;

static void orc_tsc_weights_deposit(double x, double y, double z, double mass,
                                    double xmin, double ymin, double zmin,
                                    double cellsize, orc_tscw *o) {
    double p[3];
    p[0] = (x - xmin) / cellsize;
    p[1] = (y - ymin) / cellsize;
    p[2] = (z - zmin) / cellsize;
    double pm = mass, ph = mass * 0.5;
    for (int a = 0; a < 3; a++) {
        long long nr = (long long)rint(p[a]);        /* round half even */
        float off = (float)(p[a] - (double)nr);       /* FP64 -> FP32 */
        o->near[a] = nr;
        o->sign[a] = (long long)copysignf(1.0f, off); /* +1 for +0.0 */
        float d = (float)fabs((double)off);
        float d2 = d * d;                              /* float product */
        if (a == 0) {
            float w1 = (float)((0.75 - (double)d2) * pm);
            float w3 = (float)(ph * (0.25 + (double)d * ((double)d - 1.0)));
            float w2 = (float)((double)w3 + pm * (double)d);
            o->w[a][0] = w1; o->w[a][1] = w2; o->w[a][2] = w3;
        } else {
            float w1 = (float)(0.75 - (double)d2);
            float w3 = (float)(0.5 * (0.25 + (double)d * ((double)d - 1.0)));
            float w2 = w3 + d;                         /* float sum */
            o->w[a][0] = w1; o->w[a][1] = w2; o->w[a][2] = w3;
        }
    }
}

/* Slab classification of tsc_omp2.c:210-270 for a given OpenMP team size.
 * Returns the pass (-t: lower half-slab of thread t, +t: upper half-slab). */
static int orc_tsc_mthreads(int nx, int ny, int nz, int nthreads) {
    int nxyz = nx > ny ? nx : ny;
    nxyz = nxyz > nz ? nxyz : nz;
    double xwidth = (double)nxyz / (double)nthreads;
    double hwidth = 0.5 * xwidth;
    int mthreads;
    if (hwidth < 5) mthreads = (int)((double)nxyz / (2 * 5.));
    else mthreads = nthreads;
    if (mthreads < 1) mthreads = 1;
    return mthreads;
}

/*
 * orc_tsc_deposit_f32: assign_density_TSC (tsc_omp2.c:166-506) in FP32 with
 * the reference's per-node summation order for a team of `nthreads` OpenMP
 * threads.  Per node, the reference adds all pass-A (negative target)
 * contributions in particle-index order, then all pass-B ones
 * (tsc_omp2.c:278-493); the half-slabs of one pass never touch a common
 * node, so the serial emulation below reproduces that order exactly.
 * With nthreads<=0 the build without OpenMP is emulated (every particle in
 * pass B, tsc_omp2.c:235-239).
 * Returns the number of particles whose 27-node stencil left the grid
 * (skipped here; the reference has no bounds check).
 */
long long orc_tsc_deposit_f32(const double *x, const double *y, const double *z,
                              const double *mass, const int *type, long long np,
                              int ptype, float *den, int nx, int ny, int nz,
                              double xmin, double ymin, double zmin,
                              double cellsize, int nthreads) {
    long long mx = 2 * (nx / 2 + 1);
    long long ngrid = mx * (long long)ny * nz;
    for (long long i = 0; i < ngrid; i++) den[i] = 0.0f;
    signed char *target = (signed char *)malloc((size_t)(np > 0 ? np : 1));
    int nxyz = nx > ny ? nx : ny;
    nxyz = nxyz > nz ? nxyz : nz;
    int mthreads = 1;
    for (long long i = 0; i < np; i++) target[i] = 1;
    if (nthreads > 0) {
        mthreads = orc_tsc_mthreads(nx, ny, nz, nthreads);
        double Xwidth = (double)nxyz / (double)mthreads;
        for (long long i = 0; i < np; i++) {
            if (ptype != TYPE_ALL && type[i] != ptype) continue;
            double xp;
            if (nxyz == nx) xp = (x[i] - xmin) / cellsize;
            else if (nxyz == ny) xp = (y[i] - ymin) / cellsize;
            else xp = (z[i] - zmin) / cellsize;
            int j = (int)(xp / Xwidth) + 1;
            int k = (int)((long double)(xp / Xwidth) + 0.5L) + 1;
            if (j >= mthreads + 1) target[i] = (signed char)mthreads;
            else if (k == j) target[i] = (signed char)(-j);
            else target[i] = (signed char)j;
        }
    }
    long long bad = 0;
    for (int pass = 0; pass < 2; pass++) {
        for (long long i = 0; i < np; i++) {
            if (ptype != TYPE_ALL && type[i] != ptype) continue;
            int t = target[i];
            if (pass == 0 && t >= 0) continue;
            if (pass == 1 && t <= 0) continue;
            orc_tscw w;
            orc_tsc_weights_deposit(x[i], y[i], z[i], mass[i], xmin, ymin, zmin,
                                    cellsize, &w);
            long long id[3][3];
            int ok = 1;
            for (int a = 0; a < 3; a++) {
                id[a][0] = w.near[a];
                id[a][1] = w.near[a] + w.sign[a];
                id[a][2] = w.near[a] - w.sign[a];
                /* x may reach the pad columns of the padded row, as the
                 * reference's unchecked stores do */
                long long n = a == 0 ? mx : (a == 1 ? ny : nz);
                for (int q = 0; q < 3; q++)
                    if (id[a][q] < 0 || id[a][q] >= n) ok = 0;
            }
            if (!ok) { bad++; continue; }
            /* node order as tsc_omp2.c:354-382: k outer, j middle, i inner */
            for (int c = 0; c < 3; c++)
                for (int b = 0; b < 3; b++) {
                    float wxy[3];
                    for (int a = 0; a < 3; a++) wxy[a] = w.w[0][a] * w.w[1][b];
                    for (int a = 0; a < 3; a++) {
                        long long off = id[0][a] + mx * (id[1][b] + (long long)ny * id[2][c]);
                        den[off] += wxy[a] * w.w[2][c];
                    }
                }
        }
    }
    double vf = orc_volfact();
    for (long long i = 0; i < ngrid; i++) den[i] = (float)((double)den[i] * vf);
    free(target);
    return bad;
}

/* FP64-accumulated deposit of the same FP32 contributions; output in double
 * (the exact-sum reference used to state FP32 errors). */
long long orc_tsc_deposit_f64(const double *x, const double *y, const double *z,
                              const double *mass, const int *type, long long np,
                              int ptype, double *den, int nx, int ny, int nz,
                              double xmin, double ymin, double zmin,
                              double cellsize) {
    long long mx = 2 * (nx / 2 + 1);
    long long ngrid = mx * (long long)ny * nz;
    for (long long i = 0; i < ngrid; i++) den[i] = 0.0;
    long long bad = 0;
    for (long long i = 0; i < np; i++) {
        if (ptype != TYPE_ALL && type[i] != ptype) continue;
        orc_tscw w;
        orc_tsc_weights_deposit(x[i], y[i], z[i], mass[i], xmin, ymin, zmin,
                                cellsize, &w);
        long long id[3][3];
        int ok = 1;
        for (int a = 0; a < 3; a++) {
            id[a][0] = w.near[a];
            id[a][1] = w.near[a] + w.sign[a];
            id[a][2] = w.near[a] - w.sign[a];
            long long n = a == 0 ? mx : (a == 1 ? ny : nz);
            for (int q = 0; q < 3; q++)
                if (id[a][q] < 0 || id[a][q] >= n) ok = 0;
        }
        if (!ok) { bad++; continue; }
        for (int c = 0; c < 3; c++)
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++) {
                    float wxy = w.w[0][a] * w.w[1][b];
                    float cval = wxy * w.w[2][c];
                    long long off = id[0][a] + mx * (id[1][b] + (long long)ny * id[2][c]);
                    den[off] += (double)cval;
                }
    }
    double vf = orc_volfact();
    for (long long i = 0; i < ngrid; i++) den[i] *= vf;
    return bad;
}

/* ------------------------------------------------------------------------ */
/* TSC read-back: findDen (tsc_omp2.c:40-163).                              */
/* ------------------------------------------------------------------------ */
static inline void orc_readback_axis(double p, int *nr, long *sg, float *w1,
                                     float *w2, float *w3) {
    int n = (int)rint(p);                       /* int near (tsc_omp2.c:54) */
    float off = (float)(p - (double)n);
    *nr = n;
    *sg = (long)copysignf(1.0f, off);
    float d = (float)fabs((double)off);
    float d2 = d * d;
    *w1 = (float)(0.75 - (double)d2);
    *w3 = (float)(0.5 * (0.25 + (double)d * ((double)d - 1.0)));
    *w2 = *w3 + d;
}

long long orc_findden(const double *x, const double *y, const double *z,
                      long long np, float *densph, double xmin, double ymin,
                      double zmin, double cellsize, const float *den, int nx,
                      int ny, int nz) {
    long mx = 2 * (nx / 2 + 1);
    long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long long i = 0; i < np; i++) {
        int nr[3]; long sg[3]; float w[3][3];
        double p[3] = {(x[i] - xmin) / cellsize, (y[i] - ymin) / cellsize,
                       (z[i] - zmin) / cellsize};
        for (int a = 0; a < 3; a++)
            orc_readback_axis(p[a], &nr[a], &sg[a], &w[a][0], &w[a][1], &w[a][2]);
        long id[3][3];
        int ok = 1;
        for (int a = 0; a < 3; a++) {
            id[a][0] = nr[a]; id[a][1] = nr[a] + sg[a]; id[a][2] = nr[a] - sg[a];
            long n = a == 0 ? mx : (a == 1 ? ny : nz);
            for (int q = 0; q < 3; q++)
                if (id[a][q] < 0 || id[a][q] >= n) ok = 0;
        }
        if (!ok) { bad++; densph[i] = 0.0f; continue; }
        float wxy[3][3];
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) wxy[b][a] = w[0][a] * w[1][b];
        float dsum = 0.0f;
        for (int c = 0; c < 3; c++)
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++) {
                    float g = den[id[0][a] + mx * (id[1][b] + (long)ny * id[2][c])];
                    dsum += (g * wxy[b][a]) * w[2][c];
                }
        densph[i] = dsum;
    }
    return bad;
}

/* ------------------------------------------------------------------------ */
/* Gaussian smoothing: gaussian_Smoothing (gsmooth.c:14-115), FP64.         */
/* r2c -> multiply by exp(-(pi RG_c)^2 sum (k'/n)^2) -> c2r -> /(nx ny nz). */
/* Full complex transforms of the real grid are used; the Hermitian half   */
/* the reference stores maps onto bins k and n-k, whose factors coincide.  */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } orc_cplx;

static int orc_is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* in-place FFT of one strided line; sign=-1 forward. */
static void orc_fft_line(orc_cplx *buf, orc_cplx *tmp, int n, int sign,
                         const orc_cplx *tw /* exp(-2 pi i m/n), m<n */) {
    if (orc_is_pow2(n)) {
        /* iterative radix-2, bit-reversed input */
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { orc_cplx t = buf[i]; buf[i] = buf[j]; buf[j] = t; }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int step = n / len;
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < len / 2; k++) {
                    orc_cplx w = tw[k * step];
                    if (sign > 0) w.im = -w.im;
                    orc_cplx u = buf[i + k], v = buf[i + k + len / 2];
                    orc_cplx vw = {v.re * w.re - v.im * w.im, v.re * w.im + v.im * w.re};
                    buf[i + k].re = u.re + vw.re; buf[i + k].im = u.im + vw.im;
                    buf[i + k + len / 2].re = u.re - vw.re;
                    buf[i + k + len / 2].im = u.im - vw.im;
                }
        }
    } else {
        /* direct DFT, O(n^2), exact twiddle table */
        for (int k = 0; k < n; k++) {
            double sr = 0, si = 0;
            for (int j = 0; j < n; j++) {
                long m = ((long)k * j) % n;
                double wr = tw[m].re, wi = sign > 0 ? -tw[m].im : tw[m].im;
                sr += buf[j].re * wr - buf[j].im * wi;
                si += buf[j].re * wi + buf[j].im * wr;
            }
            tmp[k].re = sr; tmp[k].im = si;
        }
        memcpy(buf, tmp, sizeof(orc_cplx) * (size_t)n);
    }
}

static orc_cplx *orc_twiddles(int n) {
    orc_cplx *tw = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)n);
    for (int m = 0; m < n; m++) {
        long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
        tw[m].re = (double)cosl(a);
        tw[m].im = (double)sinl(a);
    }
    return tw;
}

static void orc_fft3d(orc_cplx *a, int nx, int ny, int nz, int sign) {
    int nmax = nx > ny ? nx : ny;
    nmax = nmax > nz ? nmax : nz;
    orc_cplx *twx = orc_twiddles(nx), *twy = orc_twiddles(ny), *twz = orc_twiddles(nz);
#pragma omp parallel
    {
        orc_cplx *line = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)nmax);
        orc_cplx *tmp = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)nmax);
#pragma omp for schedule(static)
        for (long r = 0; r < (long)ny * nz; r++) {
            orc_cplx *row = a + r * nx;
            memcpy(line, row, sizeof(orc_cplx) * (size_t)nx);
            orc_fft_line(line, tmp, nx, sign, twx);
            memcpy(row, line, sizeof(orc_cplx) * (size_t)nx);
        }
#pragma omp for schedule(static)
        for (long r = 0; r < (long)nx * nz; r++) {
            long i = r % nx, k = r / nx;
            for (int j = 0; j < ny; j++) line[j] = a[i + (long)nx * (j + (long)ny * k)];
            orc_fft_line(line, tmp, ny, sign, twy);
            for (int j = 0; j < ny; j++) a[i + (long)nx * (j + (long)ny * k)] = line[j];
        }
#pragma omp for schedule(static)
        for (long r = 0; r < (long)nx * ny; r++) {
            for (int k = 0; k < nz; k++) line[k] = a[r + (long)nx * ny * k];
            orc_fft_line(line, tmp, nz, sign, twz);
            for (int k = 0; k < nz; k++) a[r + (long)nx * ny * k] = line[k];
        }
        free(line); free(tmp);
    }
    free(twx); free(twy); free(twz);
}

/* Per-axis factor exactly as gsmooth.c:67-92 evaluates it (double). */
static double orc_gfactor(long kk, int n, double expfac) {
    long h = n / 2;
    long k1 = kk;
    if (kk > h) k1 = kk - n; /* negative frequencies (y,z wrap; x Hermitian) */
    return exp((double)(k1 * k1) * expfac / (double)(n * n));
}

/* Expfac with the reference's float RG rescale (gsmooth.c:67-69). */
double orc_gsmooth_expfac(double cellsize, float RG) {
    float rgc = (float)((double)RG / cellsize);
    return -(M_PI * (double)rgc) * (M_PI * (double)rgc);
}

/* FP64 smoothing of the padded float grid (in place; pad columns set to 0). */
void orc_gsmooth(float *den, int nx, int ny, int nz, double cellsize, float RG) {
    long mx = 2 * (nx / 2 + 1);
    size_t nc = (size_t)nx * ny * nz;
    orc_cplx *a = (orc_cplx *)malloc(sizeof(orc_cplx) * nc);
    for (long k = 0; k < nz; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++) {
                a[i + nx * (j + ny * k)].re = den[i + mx * (j + ny * k)];
                a[i + nx * (j + ny * k)].im = 0.0;
            }
    orc_fft3d(a, nx, ny, nz, -1);
    double expfac = orc_gsmooth_expfac(cellsize, RG);
    double *fx = (double *)malloc(sizeof(double) * nx);
    double *fy = (double *)malloc(sizeof(double) * ny);
    double *fz = (double *)malloc(sizeof(double) * nz);
    for (int i = 0; i < nx; i++) fx[i] = orc_gfactor(i, nx, expfac);
    for (int j = 0; j < ny; j++) fy[j] = orc_gfactor(j, ny, expfac);
    for (int k = 0; k < nz; k++) fz[k] = orc_gfactor(k, nz, expfac);
    for (long k = 0; k < nz; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++) {
                double f = fx[i] * fy[j] * fz[k];
                a[i + nx * (j + ny * k)].re *= f;
                a[i + nx * (j + ny * k)].im *= f;
            }
    orc_fft3d(a, nx, ny, nz, +1);
    /* scale 1/(nx ny nz) formed in long double as gsmooth.c:26 */
    double scale = (double)(1.0L / (long double)((double)nx * (double)ny * (double)nz));
    for (long k = 0; k < nz; k++)
        for (long j = 0; j < ny; j++) {
            for (long i = 0; i < nx; i++)
                den[i + mx * (j + ny * k)] = (float)(a[i + nx * (j + ny * k)].re * scale);
            for (long i = nx; i < mx; i++) den[i + mx * (j + ny * k)] = 0.0f;
        }
    free(a); free(fx); free(fy); free(fz);
}

/* FP64 smoothing of a double grid (unpadded, nx*ny*nz, in place). */
void orc_gsmooth_f64(double *g, int nx, int ny, int nz, double cellsize, float RG) {
    size_t nc = (size_t)nx * ny * nz;
    orc_cplx *a = (orc_cplx *)malloc(sizeof(orc_cplx) * nc);
    for (size_t i = 0; i < nc; i++) { a[i].re = g[i]; a[i].im = 0.0; }
    orc_fft3d(a, nx, ny, nz, -1);
    double expfac = orc_gsmooth_expfac(cellsize, RG);
    for (long k = 0; k < nz; k++)
        for (long j = 0; j < ny; j++)
            for (long i = 0; i < nx; i++) {
                double f = orc_gfactor(i, nx, expfac) * orc_gfactor(j, ny, expfac) *
                           orc_gfactor(k, nz, expfac);
                a[i + nx * (j + ny * k)].re *= f;
                a[i + nx * (j + ny * k)].im *= f;
            }
    orc_fft3d(a, nx, ny, nz, +1);
    double scale = (double)(1.0L / (long double)((double)nx * (double)ny * (double)nz));
    for (size_t i = 0; i < nc; i++) g[i] = a[i].re * scale;
    free(a);
}

/* Closed-form 1-D periodic kernel: g[m] = (1/n) sum_k f(k) cos(2 pi k m / n)
 * (the delta response of one axis, written without any FFT). */
void orc_gsmooth_kernel1d(int n, double cellsize, float RG, double *g) {
    double expfac = orc_gsmooth_expfac(cellsize, RG);
    for (int m = 0; m < n; m++) {
        long double s = 0.0L;
        for (int k = 0; k < n; k++) {
            long double ang = 2.0L * 3.141592653589793238462643383279502884L *
                              (long double)(((long)k * m) % n) / (long double)n;
            s += (long double)orc_gfactor(k, n, expfac) * cosl(ang);
        }
        g[m] = (double)(s / (long double)n);
    }
}

/* ------------------------------------------------------------------------ */
/* Exact k nearest neighbours (self included), FP64 distances formed as     */
/* nnost.c:362-365 (query minus candidate, dx*dx+dy*dy+dz*dz, no FMA),      */
/* sorted ascending by (dist2, particle index).                             */
/* ------------------------------------------------------------------------ */
static inline double orc_d2(double qx, double qy, double qz, double cx,
                            double cy, double cz) {
    double tx = qx - cx, ty = qy - cy, tz = qz - cz;
    double a = tx * tx;
    double b = ty * ty;
    double c = tz * tz;
    return (a + b) + c;
}

typedef struct { double d2; long long idx; } orc_nb;

static inline int orc_nb_lt(double d2a, long long ia, double d2b, long long ib) {
    return d2a < d2b || (d2a == d2b && ia < ib);
}

/* insert into sorted list of size cnt (capacity k). */
static inline void orc_nb_insert(orc_nb *list, int *cnt, int k, double d2,
                                 long long idx) {
    int n = *cnt;
    if (n == k && !orc_nb_lt(d2, idx, list[k - 1].d2, list[k - 1].idx)) return;
    int pos = n < k ? n : k - 1;
    while (pos > 0 && orc_nb_lt(d2, idx, list[pos - 1].d2, list[pos - 1].idx)) {
        list[pos] = list[pos - 1];
        pos--;
    }
    list[pos].d2 = d2;
    list[pos].idx = idx;
    if (n < k) *cnt = n + 1;
}

/* Brute force: queries q[0..nq) vs candidates c[0..nc).  Returns 0, or -1 if
 * nc < k. */
int orc_knn_brute(const double *qx, const double *qy, const double *qz,
                  long long nq, const double *cx, const double *cy,
                  const double *cz, long long nc, int k, long long *idx,
                  double *d2out) {
    if (k <= 0 || nc < k) return -1;
#pragma omp parallel
    {
        orc_nb *list = (orc_nb *)malloc(sizeof(orc_nb) * (size_t)k);
#pragma omp for schedule(dynamic, 64)
        for (long long i = 0; i < nq; i++) {
            int cnt = 0;
            for (long long j = 0; j < nc; j++) {
                double d = orc_d2(qx[i], qy[i], qz[i], cx[j], cy[j], cz[j]);
                orc_nb_insert(list, &cnt, k, d, j);
            }
            for (int j = 0; j < k; j++) {
                idx[i * k + j] = list[j].idx;
                if (d2out) d2out[i * k + j] = list[j].d2;
            }
        }
        free(list);
    }
    return 0;
}

/* k-d tree (median split on the widest axis, leaf <= 16), exact search.
 * Node boxes are the min/max of member coordinates, so the box distance
 * (formed with the same monotone FP64 operations) never exceeds the
 * computed distance of a member. */
typedef struct {
    double lo[3], hi[3];
    long long begin, end; /* range in perm */
    int left, right;      /* -1 for leaves */
} orc_kdnode;

typedef struct {
    const double *c[3];
    long long *perm;
    orc_kdnode *nodes;
    int nnodes, cap;
} orc_kdtree;

/* (coordinate, index) order used to split k-d nodes */
static inline int orc_kd_less(const double *c, long long a, long long b) {
    return c[a] < c[b] || (c[a] == c[b] && a < b);
}

/* Quickselect: place the m-th element of perm[b,e) (in (coordinate, index)
 * order) at position m, smaller ones before it, larger ones after. */
static void orc_kd_select(long long *perm, long long b, long long e, long long m,
                          const double *c) {
    while (e - b > 1) {
        long long mid = b + (e - b) / 2;
        /* median of three as pivot */
        long long i0 = perm[b], i1 = perm[mid], i2 = perm[e - 1], piv;
        if (orc_kd_less(c, i0, i1)) {
            if (orc_kd_less(c, i1, i2)) piv = i1;
            else piv = orc_kd_less(c, i0, i2) ? i2 : i0;
        } else {
            if (orc_kd_less(c, i0, i2)) piv = i0;
            else piv = orc_kd_less(c, i1, i2) ? i2 : i1;
        }
        long long lo = b, hi = e - 1;
        while (lo <= hi) {
            while (orc_kd_less(c, perm[lo], piv)) lo++;
            while (orc_kd_less(c, piv, perm[hi])) hi--;
            if (lo <= hi) {
                long long t = perm[lo]; perm[lo] = perm[hi]; perm[hi] = t;
                lo++; hi--;
            }
        }
        if (m <= hi) e = hi + 1;
        else if (m >= lo) b = lo;
        else return;
    }
}

static int orc_kd_build(orc_kdtree *t, long long b, long long e) {
    if (t->nnodes == t->cap) {
        t->cap = t->cap * 2 + 16;
        t->nodes = (orc_kdnode *)realloc(t->nodes, sizeof(orc_kdnode) * (size_t)t->cap);
    }
    int id = t->nnodes++;
    orc_kdnode nd;
    for (int a = 0; a < 3; a++) { nd.lo[a] = INFINITY; nd.hi[a] = -INFINITY; }
    for (long long i = b; i < e; i++) {
        long long p = t->perm[i];
        for (int a = 0; a < 3; a++) {
            double v = t->c[a][p];
            if (v < nd.lo[a]) nd.lo[a] = v;
            if (v > nd.hi[a]) nd.hi[a] = v;
        }
    }
    nd.begin = b; nd.end = e; nd.left = nd.right = -1;
    if (e - b > 16) {
        int ax = 0;
        double wmax = -1;
        for (int a = 0; a < 3; a++)
            if (nd.hi[a] - nd.lo[a] > wmax) { wmax = nd.hi[a] - nd.lo[a]; ax = a; }
        long long m = b + (e - b) / 2;
        orc_kd_select(t->perm, b, e, m, t->c[ax]);
        t->nodes[id] = nd;
        int l = orc_kd_build(t, b, m);
        int r = orc_kd_build(t, m, e);
        t->nodes[id].left = l;
        t->nodes[id].right = r;
        return id;
    }
    t->nodes[id] = nd;
    return id;
}

static inline double orc_box_d2(const orc_kdnode *nd, double qx, double qy, double qz) {
    double q[3] = {qx, qy, qz};
    double g[3];
    for (int a = 0; a < 3; a++) {
        if (q[a] < nd->lo[a]) g[a] = nd->lo[a] - q[a];
        else if (q[a] > nd->hi[a]) g[a] = q[a] - nd->hi[a];
        else g[a] = 0.0;
    }
    double a0 = g[0] * g[0], a1 = g[1] * g[1], a2 = g[2] * g[2];
    return (a0 + a1) + a2;
}

static void orc_kd_search(const orc_kdtree *t, int node, double qx, double qy,
                          double qz, orc_nb *list, int *cnt, int k) {
    const orc_kdnode *nd = &t->nodes[node];
    if (*cnt == k) {
        double bd = orc_box_d2(nd, qx, qy, qz);
        if (bd > list[k - 1].d2) return;
    }
    if (nd->left < 0) {
        for (long long i = nd->begin; i < nd->end; i++) {
            long long p = t->perm[i];
            double d = orc_d2(qx, qy, qz, t->c[0][p], t->c[1][p], t->c[2][p]);
            orc_nb_insert(list, cnt, k, d, p);
        }
        return;
    }
    double dl = orc_box_d2(&t->nodes[nd->left], qx, qy, qz);
    double dr = orc_box_d2(&t->nodes[nd->right], qx, qy, qz);
    if (dl <= dr) {
        orc_kd_search(t, nd->left, qx, qy, qz, list, cnt, k);
        orc_kd_search(t, nd->right, qx, qy, qz, list, cnt, k);
    } else {
        orc_kd_search(t, nd->right, qx, qy, qz, list, cnt, k);
        orc_kd_search(t, nd->left, qx, qy, qz, list, cnt, k);
    }
}

int orc_knn_kdtree(const double *qx, const double *qy, const double *qz,
                   long long nq, const double *cx, const double *cy,
                   const double *cz, long long nc, int k, long long *idx,
                   double *d2out) {
    if (k <= 0 || nc < k) return -1;
    orc_kdtree t;
    t.c[0] = cx; t.c[1] = cy; t.c[2] = cz;
    t.perm = (long long *)malloc(sizeof(long long) * (size_t)nc);
    for (long long i = 0; i < nc; i++) t.perm[i] = i;
    t.nodes = NULL; t.nnodes = 0; t.cap = 0;
    orc_kd_build(&t, 0, nc);
#pragma omp parallel
    {
        orc_nb *list = (orc_nb *)malloc(sizeof(orc_nb) * (size_t)k);
#pragma omp for schedule(dynamic, 256)
        for (long long i = 0; i < nq; i++) {
            int cnt = 0;
            orc_kd_search(&t, 0, qx[i], qy[i], qz[i], list, &cnt, k);
            for (int j = 0; j < k; j++) {
                idx[i * k + j] = list[j].idx;
                if (d2out) d2out[i * k + j] = list[j].d2;
            }
        }
        free(list);
    }
    free(t.perm);
    free(t.nodes);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* M4 cubic spline W4 (nnost.c:534-555) and kNN densities.                  */
/* ------------------------------------------------------------------------ */
float orc_W4(float dist, float h) {
    float r = dist / h;
    if (r < 1.) {
        double t = (2. - (double)r) * (double)r * (double)r - 1.333333;
        return (float)((-0.2387324) * t / (double)(h * h * h));
    } else if (r < 2.) {
        float twomr = (float)((2. - (double)r) / (double)h);
        return (float)((0.07957747) * (double)twomr * (double)twomr * (double)twomr);
    }
    return 0.0f;
}

/* Density from sorted exact lists: rho_i = sum_j W4(r_ij, h_i) [* m_j], with
 * h_i = sqrtf(d2 of the last neighbour)/2 and r_ij = sqrtf(d2_ij), summed in
 * float in list order (nnost.c:607-623 unweighted, :689-704 weighted; the
 * neighbour mass is the FP32 copy kept in the Neighbor record, nnost.c:287). */
void orc_sph_density(const double *d2, const long long *idx, long long nq, int k,
                     const double *cmass /* nullable: unweighted */, float *dens) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < nq; i++) {
        const double *dd = d2 + i * k;
        float neardist = sqrtf((float)dd[k - 1]);
        float h = (float)((double)neardist / 2.);
        float s = 0.0f;
        for (int j = 0; j < k; j++) {
            float r = sqrtf((float)dd[j]);
            float w = orc_W4(r, h);
            if (cmass) w = w * (float)cmass[idx[i * k + j]];
            s += w;
        }
        dens[i] = s;
    }
}

/* OpenMP team size actually used (for reporting CPU baselines). */
int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ======================================================================== */
/* FP64 forward and reference gradients of the density stage (no           */
/* counterpart in the reference: the adjoint is this build's addition).     */
/* The forward here is the mathematical one in FP64 -- TSC weights of       */
/* tsc_omp2.c:303-353 / findDen tsc_omp2.c:48-161, FFT smoothing of         */
/* gsmooth.c:67-104 (orc_gsmooth_f64), W4 of nnost.c:536-555 -- so central  */
/* finite differences of it pin the analytic gradients below.  Grids are   */
/* unpadded (nz, ny, nx) doubles; every stencil must lie inside the grid.  */
/* ======================================================================== */

/* One axis: nodes (nearest rint(p), the node on the offset's side, the
 * other side), weights w(u) = 0.75 - u^2 (|u| <= 0.5), 0.5 (1.5 - |u|)^2
 * (0.5 < |u| < 1.5) for u = p - node, and dw/dp. */
static void orc_tsc_axis_d(double p, long *id, double *w, double *dw) {
    long nr = (long)rint(p);
    double off = p - (double)nr;
    long sg = off < 0.0 ? -1 : 1;
    double d = fabs(off);
    id[0] = nr; id[1] = nr + sg; id[2] = nr - sg;
    w[0] = 0.75 - d * d;            dw[0] = -2.0 * off;
    w[1] = 0.5 * (0.5 + d) * (0.5 + d); dw[1] = (0.5 + d) * (double)sg;
    w[2] = 0.5 * (0.5 - d) * (0.5 - d); dw[2] = -(0.5 - d) * (double)sg;
}

static int orc_stencil_d(double x, double y, double z, int nx, int ny, int nz,
                         double xmin, double ymin, double zmin, double h,
                         long id[3][3], double w[3][3], double dw[3][3]) {
    double p[3] = {(x - xmin) / h, (y - ymin) / h, (z - zmin) / h};
    long n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; a++) {
        orc_tsc_axis_d(p[a], id[a], w[a], dw[a]);
        for (int q = 0; q < 3; q++)
            if (id[a][q] < 0 || id[a][q] >= n[a]) return 0;
    }
    return 1;
}

/* den[(k*ny + j)*nx + i] = volfact * sum_p m_p W(x_p - node); returns the
 * number of particles whose stencil left the grid (skipped). */
long long orc_tsc_deposit_d(const double *x, const double *y, const double *z,
                            const double *m, long long np, double *den, int nx, int ny,
                            int nz, double xmin, double ymin, double zmin, double h) {
    long long ng = (long long)nx * ny * nz, bad = 0;
    for (long long i = 0; i < ng; i++) den[i] = 0.0;
    double vf = orc_volfact();
    for (long long p = 0; p < np; p++) {
        long id[3][3]; double w[3][3], dw[3][3];
        if (!orc_stencil_d(x[p], y[p], z[p], nx, ny, nz, xmin, ymin, zmin, h, id, w, dw)) {
            bad++;
            continue;
        }
        for (int c = 0; c < 3; c++)
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++)
                    den[(id[2][c] * ny + id[1][b]) * nx + id[0][a]] +=
                        vf * m[p] * w[0][a] * w[1][b] * w[2][c];
    }
    return bad;
}

/* out[p] = sum_nodes W(x_p - node) den(node). */
long long orc_readback_d(const double *x, const double *y, const double *z, long long np,
                         const double *den, int nx, int ny, int nz, double xmin,
                         double ymin, double zmin, double h, double *out) {
    long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long long p = 0; p < np; p++) {
        long id[3][3]; double w[3][3], dw[3][3];
        if (!orc_stencil_d(x[p], y[p], z[p], nx, ny, nz, xmin, ymin, zmin, h, id, w, dw)) {
            bad++;
            out[p] = 0.0;
            continue;
        }
        double s = 0.0;
        for (int c = 0; c < 3; c++)
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++)
                    s += w[0][a] * w[1][b] * w[2][c] *
                         den[(id[2][c] * ny + id[1][b]) * nx + id[0][a]];
        out[p] = s;
    }
    return bad;
}

/* Gradient of L(rho_tsc) for rho_tsc = R S D: with g = dL/drho_tsc, the
 * forward smoothed grid F = S D and the adjoint grid A = S D_g (D_g the
 * deposit of masses g, volfact included):
 *   dL/dm_p = R_p A
 *   dL/dx_p = (1/h) sum_nodes dW/dx (g_p F + m_p A)   (likewise y, z).
 * Results are added to gx, gy, gz, gm. */
void orc_tsc_grad_d(const double *x, const double *y, const double *z, const double *m,
                    const double *g, long long np, const double *F, const double *A, int nx,
                    int ny, int nz, double xmin, double ymin, double zmin, double h,
                    double *gx, double *gy, double *gz, double *gm) {
#pragma omp parallel for schedule(static)
    for (long long p = 0; p < np; p++) {
        long id[3][3]; double w[3][3], dw[3][3];
        if (!orc_stencil_d(x[p], y[p], z[p], nx, ny, nz, xmin, ymin, zmin, h, id, w, dw))
            continue;
        double sx = 0, sy = 0, sz = 0, sm = 0;
        for (int c = 0; c < 3; c++)
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++) {
                    long long o = (id[2][c] * ny + id[1][b]) * nx + id[0][a];
                    double v = g[p] * F[o] + m[p] * A[o];
                    sx += dw[0][a] * w[1][b] * w[2][c] * v;
                    sy += w[0][a] * dw[1][b] * w[2][c] * v;
                    sz += w[0][a] * w[1][b] * dw[2][c] * v;
                    sm += w[0][a] * w[1][b] * w[2][c] * A[o];
                }
        gx[p] += sx / h;
        gy[p] += sy / h;
        gz[p] += sz / h;
        gm[p] += sm;
    }
}

/* M4 spline f(q) of W(r, h) = f(r/h) / (pi h^3) and f'(q). */
static void orc_m4(double q, double *f, double *fp) {
    if (q < 1.0) {
        *f = 1.0 - 1.5 * q * q + 0.75 * q * q * q;
        *fp = -3.0 * q + 2.25 * q * q;
    } else if (q < 2.0) {
        double t = 2.0 - q;
        *f = 0.25 * t * t * t;
        *fp = -0.75 * t * t;
    } else {
        *f = 0.0;
        *fp = 0.0;
    }
}

static double orc_dist_d(const double *x, const double *y, const double *z, long long i,
                         long long j) {
    double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* rho_i = sum_t m_t W(r_it, h_i), h_i = r_i,(k-1) / 2, over the fixed lists
 * idx[i*k + t] (m == NULL: unweighted). */
void orc_sph_density_d(const double *x, const double *y, const double *z, const double *m,
                       const long long *idx, long long n, int k, double *rho) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; i++) {
        const long long *li = idx + i * k;
        double hh = 0.5 * orc_dist_d(x, y, z, i, li[k - 1]);
        double s = 0.0;
        for (int t = 0; t < k; t++) {
            double f, fp;
            orc_m4(orc_dist_d(x, y, z, i, li[t]) / hh, &f, &fp);
            s += (m ? m[li[t]] : 1.0) * f / (M_PI * hh * hh * hh);
        }
        rho[i] = s;
    }
}

/* Gradient of L(rho_knn) over fixed lists, g = dL/drho_knn:
 *   dW/dr = f'(q) / (pi h^4),  dW/dh = -(3 f(q) + q f'(q)) / (pi h^4),
 *   drho_i = sum_t [m_t dW/dr dr_it + m_t dW/dh dh_i + W dm_t],
 *   dr_it = (x_i - x_t).(dx_i - dx_t) / r_it,  dh_i = dr_i,(k-1) / 2.
 * Serial (neighbours receive contributions).  Added to gx, gy, gz, gm. */
void orc_sph_grad_d(const double *x, const double *y, const double *z, const double *m,
                    const long long *idx, long long n, int k, const double *g, double *gx,
                    double *gy, double *gz, double *gm) {
    for (long long i = 0; i < n; i++) {
        const long long *li = idx + i * k;
        const long long K = li[k - 1];
        double rK = orc_dist_d(x, y, z, i, K), hh = 0.5 * rK;
        double h4 = M_PI * hh * hh * hh * hh;
        double sh = 0.0;
        for (int t = 0; t < k; t++) {
            long long j = li[t];
            double r = orc_dist_d(x, y, z, i, j), q = r / hh, f, fp;
            orc_m4(q, &f, &fp);
            double mt = m ? m[j] : 1.0;
            sh += mt * (-(3.0 * f + q * fp) / h4);
            if (m) gm[j] += g[i] * f / (M_PI * hh * hh * hh);
            if (r > 0.0) {
                double c = g[i] * mt * (fp / h4) / r;
                double ex = c * (x[i] - x[j]), ey = c * (y[i] - y[j]), ez = c * (z[i] - z[j]);
                gx[i] += ex; gy[i] += ey; gz[i] += ez;
                gx[j] -= ex; gy[j] -= ey; gz[j] -= ez;
            }
        }
        if (rK > 0.0) {
            double c = g[i] * sh * 0.5 / rK;
            double ex = c * (x[i] - x[K]), ey = c * (y[i] - y[K]), ez = c * (z[i] - z[K]);
            gx[i] += ex; gy[i] += ey; gz[i] += ez;
            gx[K] -= ex; gy[K] -= ey; gz[K] -= ez;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Integer peak stage.                                                      */
/* ------------------------------------------------------------------------ */
#define ORC_TYPE_STAR 1

/* Grid cores of lagFindStellarCore (nnost.c:1115-1271).  Particles (stars
 * for ptype 1, all for ptype 5, none otherwise) are chained per nearest grid
 * node (rint of the offset in cells) by head insertion in index order, so a
 * node's chain runs from its largest index down (nnost.c:1121-1156).  A node
 * of the interior [nbuff/2, n - nbuff/2) on every axis whose smoothed value
 * exceeds `threshold` and every one of its 26 neighbours, and whose chain is
 * not empty, gives a core: the first chain member of maximal densph
 * (strictly greater replaces, nnost.c:1243-1249) -- unless the chain's last
 * member, its smallest index, is particle 0 (`if(ibp > 0)`, nnost.c:1202,
 * 1254).  Cores in (k, j, i) node order, as the OpenMP static schedule over k
 * concatenates them (nnost.c:1216-1266).  Writes up to maxcore core
 * particles; returns the number of cores. */
long long orc_grid_cores(const double *x, const double *y, const double *z, const int *type,
                         long long np, int ptype, const float *densph, const float *den,
                         int nx, int ny, int nz, double xmin, double ymin, double zmin,
                         double cellsize, double threshold, int nbuff, long long *core_peak,
                         long long maxcore) {
    const long long mx = 2LL * (nx / 2 + 1);
    const long long ncells = mx * (long long)ny * nz;
    long long *head = (long long *)malloc(sizeof(long long) * (size_t)ncells);
    long long *next = (long long *)malloc(sizeof(long long) * (size_t)(np > 0 ? np : 1));
    for (long long c = 0; c < ncells; c++) head[c] = -1;
    if (ptype == ORC_TYPE_STAR || ptype == TYPE_ALL) {
        for (long long i = 0; i < np; i++) {
            if (ptype == ORC_TYPE_STAR && type[i] != ORC_TYPE_STAR) continue;
            const long long a = (long long)rint((x[i] - xmin) / cellsize);
            const long long b = (long long)rint((y[i] - ymin) / cellsize);
            const long long c = (long long)rint((z[i] - zmin) / cellsize);
            const long long cell = a + mx * (b + ny * c);
            next[i] = head[cell];
            head[cell] = i;
        }
    }
    long long ncore = 0;
    const int h = nbuff / 2;
    for (int k = h; k < nz - h; k++)
        for (int j = h; j < ny - h; j++)
            for (int i = h; i < nx - h; i++) {
                const long long cell = i + mx * ((long long)j + (long long)ny * k);
                const float v = den[cell];
                if (!((double)v > threshold) || head[cell] < 0) continue;
                int peak = 1;
                for (int dk = -1; dk <= 1 && peak; dk++)
                    for (int dj = -1; dj <= 1 && peak; dj++)
                        for (int di = -1; di <= 1 && peak; di++) {
                            if (!di && !dj && !dk) continue;
                            const long long o = cell + di + mx * (dj + (long long)ny * dk);
                            if (!(v > den[o])) peak = 0;
                        }
                if (!peak) continue;
                float best = -1.e20f;
                long long arg = -1, last = -1;
                for (long long p = head[cell]; p >= 0; p = next[p]) {
                    if (densph[p] > best) {
                        best = densph[p];
                        arg = p;
                    }
                    last = p;
                }
                if (last > 0) {
                    if (ncore < maxcore) core_peak[ncore] = arg;
                    ncore++;
                }
            }
    free(head);
    free(next);
    return ncore;
}

/* finddenpeak (subhaloden.mod6.c:416-492): particle i is a core when its
 * density exceeds `threshold` (PEAKTHRESHOLD in the reference), it is a star
 * if jflag == 1, and no particle of its neighbour list has a strictly larger
 * density.  Cores in index order; writes up to maxcore; returns the count. */
long long orc_finddenpeak(const float *den, int k, const long long *nbr, long long np,
                          const int *type, int jflag, double threshold, long long *core_peak,
                          long long maxcore) {
    long long ncore = 0;
    for (long long i = 0; i < np; i++) {
        if (!((double)den[i] > threshold)) continue;
        if (jflag == 1 && type[i] != ORC_TYPE_STAR) continue;
        int peak = 1;
        for (int t = 0; t < k && peak; t++)
            if (den[nbr[i * k + t]] > den[i]) peak = 0;
        if (peak) {
            if (ncore < maxcore) core_peak[ncore] = i;
            ncore++;
        }
    }
    return ncore;
}

/* ------------------------------------------------------------------------ */
/* MergingPeak (subhaloden.mod6.c:341-415; friends-of-friends walk of       */
/* ost.c:306-356).  Cores are ordered by key, largest first (the qsort with */
/* sortcoreden, :323-330; equal keys keep their input order, as glibc's     */
/* merge sort does).  In that order, each core not yet claimed claims every */
/* unclaimed core reachable through links and stays a core itself; two      */
/* cores link when (double)sqrtf((float)d2) <= link, d2 the FP64 squared   */
/* distance of their peak particles formed dx*dx+dy*dy+dz*dz (ost.c:       */
/* 328-333).  x, y, z: each core's peak position.  order[r]: the input      */
/* index of the core of rank r.  keep[0 .. return): the kept cores' input   */
/* indices in rank order.  A breadth-first claim over all unclaimed cores  */
/* per step: O(n^2), for test sizes.                                        */
/* ------------------------------------------------------------------------ */
static void orc_stable_desc(const float *key, int *idx, int *tmp, long long n) {
    for (long long w = 1; w < n; w *= 2) {
        for (long long lo = 0; lo < n; lo += 2 * w) {
            long long mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            long long a = lo, b = mid, t = lo;
            while (a < mid && b < hi) tmp[t++] = key[idx[b]] > key[idx[a]] ? idx[b++] : idx[a++];
            while (a < mid) tmp[t++] = idx[a++];
            while (b < hi) tmp[t++] = idx[b++];
        }
        memcpy(idx, tmp, sizeof(int) * (size_t)n);
    }
}

static int orc_fof_linked(const double *x, const double *y, const double *z, int a, int b,
                          double link) {
    double dx = x[a] - x[b], dy = y[a] - y[b], dz = z[a] - z[b];
    double d2 = dx * dx + dy * dy + dz * dz;
    float d = sqrtf((float)d2);
    return (double)d <= link;
}

long long orc_merging_peak(const double *x, const double *y, const double *z, const float *key,
                           long long n, double link, int *order, int *keep) {
    if (n <= 0) return 0;
    int *tmp = (int *)malloc(sizeof(int) * (size_t)n);
    int *queue = (int *)malloc(sizeof(int) * (size_t)n);
    unsigned char *claimed = (unsigned char *)calloc((size_t)n, 1);
    for (long long i = 0; i < n; i++) order[i] = (int)i;
    orc_stable_desc(key, order, tmp, n);
    long long nkeep = 0;
    for (long long r = 0; r < n; r++) {
        const int s = order[r];
        if (claimed[s]) continue;
        long long head = 0, tail = 0;
        claimed[s] = 1;
        queue[tail++] = s;
        while (head < tail) {
            const int a = queue[head++];
            for (long long b = 0; b < n; b++)
                if (!claimed[b] && orc_fof_linked(x, y, z, a, (int)b, link)) {
                    claimed[b] = 1;
                    queue[tail++] = (int)b;
                }
        }
        keep[nkeep++] = s;
    }
    free(tmp);
    free(queue);
    free(claimed);
    return nkeep;
}

/* ------------------------------------------------------------------------ */
/* Watershed of subhalo_den: FindCoreDensity, the shell thresholds and      */
/* GetShellParticlesPeaks + SaveRemainingParticles2LastShell.               */
/* Working-particle flag bits (subhaloden.mod6.c:180-186) and the          */
/* non-member id (defs.h:29).                                               */
/* ------------------------------------------------------------------------ */
#define ORC_WP_PEAK 0x01
#define ORC_WP_VISITED 0x02
#define ORC_WP_CORE 0x04
#define ORC_WP_SHELL 0x08
#define ORC_WP_BOUND 0x10
#define ORC_WP_REMAINING 0x20
#define ORC_NOT_HALO_MEMBER (-9999)
#define ORC_CORE_ENCLOSED 0x01

/* One breadth-first sweep of FindCoreDensity from `root` over particles
 * denser than thr (subhaloden.mod6.c:1243-1266 for the trial sweep,
 * :1276-1300 for the final one).  A neighbour that is an unvisited peak
 * sets the break flag; the trial sweep (stop_all) then ends at once, the
 * final sweep only leaves the rest of that particle's list.  vis[] is the
 * thread's visited bit; the pushed particles are q[0 .. return). */
static long long orc_core_sweep(const float *den, const long long *nbr, int k,
                                const unsigned char *flag, unsigned char *vis, int *q,
                                int root, float thr, int stop_all, int *brk) {
    long long n = 0, now = 0;
    *brk = 0;
    q[n++] = root;
    vis[root] = 1;
    while (now < n && !(stop_all && *brk)) {
        const long long base = (long long)q[now] * k;
        for (int s = 0; s < k; s++) {
            const long long v = nbr[base + s];
            if (!(den[v] > thr)) continue;
            if (!vis[v] && !(flag[v] & ORC_WP_PEAK)) {
                vis[v] = 1;
                q[n++] = (int)v;
            } else if (!vis[v]) {
                *brk = 1;
                break;
            }
        }
        now++;
    }
    return n;
}

/* FindCoreDensity (subhaloden.mod6.c:1167-1475) as one OpenMP thread runs
 * it (the reference parallelises over cores with per-thread visited bits;
 * with one thread its result is the schedule-free one).  den: wp[].den as
 * Dump2WorkingParticle sets it (:493-500); peak[c]: core[c].peak.
 * Per core: coredensity, nummem, numstar, starmass (the FP32 sum of the
 * stars' FP64 masses in sweep order); per particle: flag (PEAK, CORE,
 * BOUND) and haloid (the last core whose final sweep holds it).
 * The threshold search (:1229-1269) halves [downden, upden] from
 * [minden, den[peak]] while |(upden-downden)/denthr| > 1e-3
 * (COREDENRESOLUTION, params.h:12; the definition at
 * subhaloden.mod6.c:22 is commented out).  A final sweep that breaks discards the core's
 * members (nummem 0) without clearing their visited bits (:1301-1310), so
 * they stay visited for every later core: restated here with one vis[]
 * kept across cores. */
long long orc_find_core_density(const float *den, long long np, const long long *nbr, int k,
                                const int *type, const double *mass, const int *peak,
                                int numcore, float *coredensity, int *nummem, int *numstar,
                                float *starmass, int *haloid, unsigned char *flag) {
    for (long long i = 0; i < np; i++) haloid[i] = ORC_NOT_HALO_MEMBER;
    float minden = 1.e23f;
    for (long long i = 0; i < np; i++) minden = den[i] < minden ? den[i] : minden;
    minden = (float)(1.0 > (double)minden ? 1.0 : (double)minden); /* DENFLOOR, params.h:109 */
    unsigned char *vis = (unsigned char *)calloc((size_t)(np > 0 ? np : 1), 1);
    int *q = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    for (long long i = 0; i < np; i++) flag[i] = 0;
    for (int c = 0; c < numcore; c++) flag[peak[c]] |= ORC_WP_PEAK;
    for (int c = 0; c < numcore; c++) {
        float upden = den[peak[c]], downden = minden, denthr;
        long long n = 0;
        int brk;
        do {
            for (long long j = 0; j < n; j++) vis[q[j]] = 0;
            denthr = (float)(0.5 * (double)(upden + downden));
            n = orc_core_sweep(den, nbr, k, flag, vis, q, peak[c], denthr, 1, &brk);
            if (!brk) upden = denthr;
            else downden = denthr;
        } while (fabs((double)((upden - downden) / denthr)) > 1.e-3);
        coredensity[c] = denthr = upden;
        for (long long j = 0; j < n; j++) vis[q[j]] = 0;
        n = orc_core_sweep(den, nbr, k, flag, vis, q, peak[c], denthr, 0, &brk);
        if (brk) n = 0;
        for (long long j = 0; j < n; j++) vis[q[j]] = 0;
        int nstar = 0;
        float smass = 0.0f;
        for (long long j = 0; j < n; j++)
            if (type[q[j]] == ORC_TYPE_STAR) {
                nstar++;
                smass = (float)((double)smass + mass[q[j]]);
            }
        for (long long j = 0; j < n; j++) {
            flag[q[j]] = (unsigned char)((flag[q[j]] | ORC_WP_CORE | ORC_WP_BOUND) &
                                         ~ORC_WP_REMAINING);
            haloid[q[j]] = c;
        }
        nummem[c] = (int)n;
        numstar[c] = nstar;
        starmass[c] = smass;
    }
    free(vis);
    free(q);
    return numcore;
}

/* The shell thresholds of subhalo_den (subhaloden.mod6.c:2995-3024): the
 * log10(1 + density) range from the floored particle minimum to the largest
 * core density, divided into nshelldivide steps (10, or
 * min(15, 10 log10(1.5 np)) when np > 1000; NSHELLDIVIDE params.h:43),
 * highest first.  Writes thr[0 .. return). */
int orc_shell_thresholds(const float *den, long long np, const float *coredensity, int numcore,
                         float *thr) {
    float minshellden = 1.e27f, maxshellden = -1.e27f;
    for (int c = 0; c < numcore; c++)
        maxshellden = maxshellden > coredensity[c] ? maxshellden : coredensity[c];
    for (long long i = 0; i < np; i++) minshellden = minshellden < den[i] ? minshellden : den[i];
    minshellden = (float)(1.0 > (double)minshellden ? 1.0 : (double)minshellden);
    minshellden = (float)log10((double)minshellden + 1.0);
    maxshellden = (float)log10((double)maxshellden + 1.0);
    int nsd = 10;
    if (np > 1000) {
        double a = 10.0 * log10((double)np * 1.5);
        nsd = (int)(10.0 > a ? 10.0 : a);
        nsd = nsd < 15 ? nsd : 15;
    }
    for (int i = 0; i < nsd; i++) {
        float t = minshellden + (float)(nsd - 1 - i) * (maxshellden - minshellden) / (float)nsd;
        thr[i] = (float)(pow(10.0, (double)t) - 1.0);
    }
    return nsd;
}

/* GetShellParticlesPeaks (subhaloden.mod6.c:612-678) once per threshold,
 * then SaveRemainingParticles2LastShell (:572-611).  For each threshold, in
 * core order, every core above it that no earlier sweep of this threshold
 * reached roots a sweep over particles denser than the threshold (peaks
 * included, no barrier); the sweep's particles that are neither shell nor
 * core particles yet form a new shell (flag SHELL, haloid non-member), and
 * the cores whose peak it reached (the first core of each such peak) are
 * the shell's cores.  Particles left outside every core and shell join the
 * last shell (or a shell of all cores when there is none).  Shells go to
 * CSR arrays: particles part[part_ptr[s] .. part_ptr[s+1]), cores
 * cores[core_ptr[s] .. core_ptr[s+1]).  flag/haloid are updated in place
 * (as FindCoreDensity left them).  A core reached by several sweeps of one
 * threshold is listed in each of their shells, as the reference does.
 * Returns nshell, -1 when more than maxshell shells, -2 when a root's own
 * core is never enclosed (two cores sharing a peak: the reference loops for
 * ever), -3 when the shell cores exceed maxcores. */
long long orc_core_shells(const float *den, long long np, const long long *nbr, int k,
                          const int *peak, const float *coredensity, int numcore,
                          const float *thr, int nthr, unsigned char *flag, int *haloid,
                          long long *part_ptr, int *part, long long *core_ptr, int *cores,
                          long long maxshell, long long maxcores) {
    int *q = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    unsigned char *cflag = (unsigned char *)calloc((size_t)(numcore > 0 ? numcore : 1), 1);
    int *first_core = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    for (long long i = 0; i < np; i++) first_core[i] = -1;
    for (int c = numcore - 1; c >= 0; c--) first_core[peak[c]] = c;
    long long nshell = 0, npart = 0, ncores = 0, n = 0;
    part_ptr[0] = core_ptr[0] = 0;
    for (int t = 0; t < nthr; t++) {
        const float dthr = thr[t];
        memset(cflag, 0, (size_t)(numcore > 0 ? numcore : 1));
        for (long long i = 0; i < np; i++) flag[i] &= (unsigned char)~ORC_WP_VISITED;
        n = 0;
        /* GET_NONENCLOSED_ID (:555-563) scans from core 0 each time; the
         * cores before the last root are enclosed or below the threshold
         * and stay so within one threshold, so the scan resumes there */
        int cursor = 0;
        for (;;) {
            int root = -1;
            for (int c = cursor; c < numcore; c++)
                if (!(cflag[c] & ORC_CORE_ENCLOSED) && coredensity[c] > dthr) {
                    root = c;
                    break;
                }
            if (root < 0) break;
            cursor = root;
            for (long long j = 0; j < n; j++) flag[q[j]] &= (unsigned char)~ORC_WP_VISITED;
            long long now = 0;
            n = 0;
            q[n++] = peak[root];
            flag[peak[root]] |= ORC_WP_VISITED;
            while (now < n) {
                const long long base = (long long)q[now] * k;
                for (int s = 0; s < k; s++) {
                    const long long v = nbr[base + s];
                    if (den[v] > dthr && !(flag[v] & ORC_WP_VISITED)) {
                        flag[v] |= ORC_WP_VISITED;
                        q[n++] = (int)v;
                    }
                }
                now++;
            }
            if (nshell >= maxshell) {
                nshell = -1;
                goto done;
            }
            for (long long j = 0; j < n; j++) {
                const int v = q[j];
                if (!(flag[v] & ORC_WP_SHELL) && !(flag[v] & ORC_WP_CORE)) {
                    part[npart++] = v;
                    flag[v] = (unsigned char)((flag[v] | ORC_WP_SHELL) & ~ORC_WP_REMAINING);
                    haloid[v] = ORC_NOT_HALO_MEMBER;
                }
            }
            for (long long j = 0; j < n; j++) {
                const int v = q[j];
                if (!(flag[v] & ORC_WP_PEAK)) continue;
                /* the reference scans the cores for the first with this peak
                 * (:664-672); first_core[] holds that core */
                const int c = first_core[v];
                if (c < 0) continue;
                if (ncores >= maxcores) {
                    nshell = -3;
                    goto done;
                }
                cflag[c] |= ORC_CORE_ENCLOSED;
                cores[ncores++] = c;
            }
            if (!(cflag[root] & ORC_CORE_ENCLOSED)) {
                nshell = -2;
                goto done;
            }
            nshell++;
            part_ptr[nshell] = npart;
            core_ptr[nshell] = ncores;
        }
    }
    {
        long long nrest = 0;
        for (long long i = 0; i < np; i++)
            if (!(flag[i] & ORC_WP_CORE) && !(flag[i] & ORC_WP_SHELL)) nrest++;
        if (nrest > 0) {
            if (nshell == 0) {
                if (maxshell < 1) {
                    nshell = -1;
                    goto done;
                }
                if (ncores + numcore > maxcores) {
                    nshell = -3;
                    goto done;
                }
                for (int c = 0; c < numcore; c++) cores[ncores++] = c;
                nshell = 1;
                core_ptr[1] = ncores;
            }
            for (long long i = 0; i < np; i++)
                if (!(flag[i] & ORC_WP_CORE) && !(flag[i] & ORC_WP_SHELL)) {
                    flag[i] = (unsigned char)((flag[i] | ORC_WP_SHELL) & ~ORC_WP_REMAINING);
                    haloid[i] = ORC_NOT_HALO_MEMBER;
                    part[npart++] = (int)i;
                }
            part_ptr[nshell] = npart;
        }
    }
done:
    free(q);
    free(cflag);
    free(first_core);
    return nshell;
}
