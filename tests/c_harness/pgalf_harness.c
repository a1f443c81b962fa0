/*
 * pgalf_harness.c -- a C caller of the density-stage library, written the
 * way NewGalFinder calls the stage: it includes pgalf_density.h, links
 * -lpgalf_density and calls the reference entry points with host arrays
 * (assign_density_TSC, gaussian_Smoothing, findDen, findsphdensity,
 * starfindsphdensity; tsc_omp2.c:40,166, gsmooth.c:14, nnost.c:560,632).
 *
 *   pgalf_harness DIR
 *
 * DIR/params.txt    "np nx ny nz xmin ymin zmin cellsize RG k ptype"
 * DIR/particles.bin np 88-byte particle records (SimpleBasicParticleType)
 * DIR/grid_in.bin   a padded float grid, mx*ny*nz (mx = 2(nx/2+1))
 * Writes DIR/tsc.bin (deposit of the particles), DIR/findden.bin (read-back
 * of grid_in), DIR/smooth.bin (grid_in smoothed), DIR/near.bin +
 * DIR/sphden.bin (findsphdensity), DIR/star_near.bin + DIR/star_den.bin
 * (starfindsphdensity), and prints each call's status.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pgalf_density.h"

static void *load(const char *dir, const char *name, size_t bytes) {
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "rb");
    if (!f) { perror(path); exit(2); }
    void *p = malloc(bytes ? bytes : 1);
    if (fread(p, 1, bytes, f) != bytes) { fprintf(stderr, "%s: short read\n", path); exit(2); }
    fclose(f);
    return p;
}

static void save(const char *dir, const char *name, const void *p, size_t bytes) {
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(p, 1, bytes, f) != bytes) { perror(path); exit(2); }
    fclose(f);
}

static int status(const char *what) {
    int e = pgalf_last_error();
    printf("%s: %d (%s)\n", what, e, pgalf_last_error_string());
    return e;
}

int main(int argc, char **argv) {
    if (argc != 2) { fprintf(stderr, "usage: %s DIR\n", argv[0]); return 2; }
    const char *dir = argv[1];
    char path[4096];
    snprintf(path, sizeof(path), "%s/params.txt", dir);
    FILE *f = fopen(path, "r");
    if (!f) { perror(path); return 2; }
    int np, nx, ny, nz, k, ptype;
    double xmin, ymin, zmin, cellsize;
    float RG;
    if (fscanf(f, "%d %d %d %d %lf %lf %lf %lf %f %d %d", &np, &nx, &ny, &nz, &xmin, &ymin,
               &zmin, &cellsize, &RG, &k, &ptype) != 11) {
        fprintf(stderr, "bad params\n");
        return 2;
    }
    fclose(f);
    const size_t ncells = (size_t)(2 * (nx / 2 + 1)) * ny * nz;
    pgalf_particle_t *bp = load(dir, "particles.bin", sizeof(pgalf_particle_t) * (size_t)np);
    float *grid_in = load(dir, "grid_in.bin", sizeof(float) * ncells);
    int bad = 0;

    float *den = calloc(ncells, sizeof(float));
    assign_density_TSC(bp, np, den, nx, ny, nz, xmin, ymin, zmin, cellsize, ptype);
    bad |= status("assign_density_TSC");
    save(dir, "tsc.bin", den, sizeof(float) * ncells);

    float *densph = calloc((size_t)np, sizeof(float));
    findDen(bp, np, densph, xmin, ymin, zmin, cellsize, grid_in, nx, ny, nz);
    bad |= status("findDen");
    save(dir, "findden.bin", densph, sizeof(float) * (size_t)np);

    memcpy(den, grid_in, sizeof(float) * ncells);
    gaussian_Smoothing(den, nx, ny, nz, cellsize, RG);
    bad |= status("gaussian_Smoothing");
    save(dir, "smooth.bin", den, sizeof(float) * ncells);

    long long *near = calloc((size_t)np * k, sizeof(long long));
    findsphdensity(bp, np, near, k, densph);
    bad |= status("findsphdensity");
    save(dir, "near.bin", near, sizeof(long long) * (size_t)np * k);
    save(dir, "sphden.bin", densph, sizeof(float) * (size_t)np);

    starfindsphdensity(bp, np, near, k, densph);
    bad |= status("starfindsphdensity");
    save(dir, "star_near.bin", near, sizeof(long long) * (size_t)np * k);
    save(dir, "star_den.bin", densph, sizeof(float) * (size_t)np);

    free(bp); free(grid_in); free(den); free(densph); free(near);
    return bad ? 1 : 0;
}
