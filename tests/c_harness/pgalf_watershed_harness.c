/*
 * pgalf_watershed_harness.c -- a C caller of the library's watershed,
 * written the way subhalo_den runs it after the peaks
 * (NewGalFinder/subhaloden.mod6.c:2978-3024): FindCoreDensity, the shell
 * thresholds, GetShellParticlesPeaks for each of them and
 * SaveRemainingParticles2LastShell, with the working particles `wp` as three
 * arrays and the shells as CSR arrays.
 *
 *   pgalf_watershed_harness DIR
 *
 * DIR/params.txt    "np k numcore"
 * DIR/particles.bin np 88-byte particle records (type, mass used)
 * DIR/neighbor.bin  np*k int64 neighbour lists
 * DIR/den.bin       np float densities (wp[].den)
 * DIR/peak.bin      numcore int core peak particles
 * Writes DIR/core.bin (numcore 56-byte core records), DIR/haloid.bin,
 * DIR/flag.bin (after the shells), DIR/thr.bin, DIR/shells.bin (nshell,
 * part_ptr[nshell+1], part[...], core_ptr[nshell+1], cores[...] as int64 /
 * int32 arrays in that order), and prints each call's status.
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

static FILE *create(const char *dir, const char *name) {
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "wb");
    if (!f) { perror(path); exit(2); }
    return f;
}

static void save(const char *dir, const char *name, const void *p, size_t bytes) {
    FILE *f = create(dir, name);
    if (fwrite(p, 1, bytes, f) != bytes) { perror(name); exit(2); }
    fclose(f);
}

int main(int argc, char **argv) {
    if (argc != 2) { fprintf(stderr, "usage: %s DIR\n", argv[0]); return 2; }
    const char *dir = argv[1];
    char path[4096];
    snprintf(path, sizeof(path), "%s/params.txt", dir);
    FILE *f = fopen(path, "r");
    if (!f) { perror(path); return 2; }
    int np, k, numcore;
    if (fscanf(f, "%d %d %d", &np, &k, &numcore) != 3) { fprintf(stderr, "bad params\n"); return 2; }
    fclose(f);
    pgalf_particle_t *bp = load(dir, "particles.bin", sizeof(pgalf_particle_t) * (size_t)np);
    long long *neighbor = load(dir, "neighbor.bin", sizeof(long long) * (size_t)np * k);
    float *den = load(dir, "den.bin", sizeof(float) * (size_t)np);
    int *peak = load(dir, "peak.bin", sizeof(int) * (size_t)numcore);
    pgalf_core_t *core = calloc((size_t)numcore, sizeof(pgalf_core_t));
    for (int c = 0; c < numcore; c++) core[c].peak = peak[c];
    int *haloid = malloc(sizeof(int) * (size_t)np);
    unsigned char *flag = malloc((size_t)np);

    /* subhaloden.mod6.c:2981 */
    int n = pgalf_find_core_density(bp, np, neighbor, k, core, numcore, den, haloid, flag);
    printf("pgalf_find_core_density: %d (%s)\n", n, pgalf_last_error_string());
    if (n != numcore) return 1;
    /* :2995-3020 */
    float thr[PGALF_MAX_SHELL_DIVIDE];
    int nthr = pgalf_shell_thresholds(den, np, core, numcore, thr);
    printf("pgalf_shell_thresholds: %d (%s)\n", nthr, pgalf_last_error_string());
    if (nthr < 1) return 1;
    /* :3021-3025 */
    pgalf_shells_t sh;
    sh.maxshell = (long long)numcore * nthr + 1;
    sh.maxcores = 64LL * (numcore + 1) * (nthr + 1);
    sh.part_ptr = malloc(sizeof(long long) * (size_t)(sh.maxshell + 1));
    sh.core_ptr = malloc(sizeof(long long) * (size_t)(sh.maxshell + 1));
    sh.part = malloc(sizeof(int) * (size_t)np);
    sh.cores = malloc(sizeof(int) * (size_t)sh.maxcores);
    sh.nshell = 0;
    int e = pgalf_shell_particles_peaks(np, neighbor, k, core, numcore, den, haloid, flag, thr,
                                        nthr, &sh);
    printf("pgalf_shell_particles_peaks: %d (%s), %lld shells\n", e, pgalf_last_error_string(),
           sh.nshell);
    if (e != PGALF_OK) return 1;

    save(dir, "core.bin", core, sizeof(pgalf_core_t) * (size_t)numcore);
    save(dir, "haloid.bin", haloid, sizeof(int) * (size_t)np);
    save(dir, "flag.bin", flag, (size_t)np);
    save(dir, "thr.bin", thr, sizeof(float) * (size_t)nthr);
    FILE *o = create(dir, "shells.bin");
    long long ns = sh.nshell;
    fwrite(&ns, sizeof(ns), 1, o);
    fwrite(sh.part_ptr, sizeof(long long), (size_t)ns + 1, o);
    fwrite(sh.part, sizeof(int), (size_t)sh.part_ptr[ns], o);
    fwrite(sh.core_ptr, sizeof(long long), (size_t)ns + 1, o);
    fwrite(sh.cores, sizeof(int), (size_t)sh.core_ptr[ns], o);
    fclose(o);
    return 0;
}
