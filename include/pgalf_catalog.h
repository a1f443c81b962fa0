/*
 * pgalf_catalog.h -- FoF catalogue loader and GALFIND catalogue writer of
 * libpgalf_density.so (host code; no GPU needed).
 *
 * Loader: reads NewGalFinder's input catalogue (gfind.c:25-91,221-241,
 * 363-371: FoF_halo_cat.* = 7-float header + HaloQ records, and
 * FoF_member_particle.* = each halo's DM, gas, sink and star records) in
 * batches of halos, localises every halo as subhalo_den does before the
 * density stage (Dump2BP, subhaloden.mod6.c:2810-2837, then mklocalize,
 * :79-177) and sizes its TSC grid as lagFindStellarCore does
 * (nnost.c:1040-1058), ready for pgalf_batch_create.
 *
 * Writer: write_data of gfind.c:700-732 (GALFIND.DATA, GALCATALOG.LIST and
 * the background particle file), byte for byte.
 *
 * Record layouts are those of the reference build (ramses.h, tree.h,
 * defs.h with -DNCHEM=3 -DXYZDBL): DmType = StarType 136 bytes, GasType
 * 112, SinkType 168, HaloQ 96, HaloInfo 112, SubInfo 112.
 */
#ifndef PGALF_CATALOG_H
#define PGALF_CATALOG_H

#include "pgalf_density.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The catalogue header (gfind.c:225-231). */
typedef struct pgalf_cat_header {
    float size, hubble, omep, omepb, omeplam, amax, anow;
} pgalf_cat_header_t;

/* HaloQ (tree.h:55-60, positions double under -DXYZDBL). */
typedef struct pgalf_haloq {
    unsigned long long np, npstar, npgas, npdm, npsink;
    double x, y, z;
    float mass, mstar, mgas, mdm, msink;
    float vx, vy, vz;
} pgalf_haloq_t;

/* HaloInfo and SubInfo (defs.h:3-27). */
typedef struct pgalf_halo_info {
    int nsub, ndm, nstar, nsink, ngas, npall;
    double totm, mdm, mgas, msink, mstar;
    double x, y, z, vx, vy, vz;
} pgalf_halo_info_t;
typedef struct pgalf_sub_info {
    int npdm, npgas, npsink, npstar, npall;
    double totm, mdm, mgas, msink, mstar;
    double x, y, z, vx, vy, vz;
} pgalf_sub_info_t;

/* Unwrapping of halos that span the periodic box (mklocalize): the
 * reference's gap search "sorts" with a comparator of element addresses
 * (subhaloden.mod6.c:57-64), which under glibc's merge sort leaves the
 * coordinates in file order; PGALF_UNWRAP_REFERENCE reproduces that,
 * PGALF_UNWRAP_SORTED uses the widest gap of the sorted coordinates. */
#define PGALF_UNWRAP_REFERENCE 0
#define PGALF_UNWRAP_SORTED 1
#define PGALF_MIN_HALO_NP 30 /* gfind.c:388 */

typedef struct pgalf_catalog pgalf_catalog_t;
/* Opens the two files and reads the header; NULL on error. */
pgalf_catalog_t *pgalf_catalog_open(const char *halo_cat, const char *members, int unwrap);
int pgalf_catalog_header(const pgalf_catalog_t *cat, pgalf_cat_header_t *out);
/* Reads the next batch: up to max_halos halos of at least min_np particles
 * (smaller ones are passed by, as gfind.c:388 does), stopping before the
 * particle total would exceed max_particles (a larger halo is read alone).
 * Returns the halo count (0 at the end, -1 on error); *np_out = particles. */
long long pgalf_catalog_next(pgalf_catalog_t *cat, long long max_halos, long long max_particles,
                             int min_np, long long *np_out);
/* Copies out the last batch (any pointer may be NULL): per particle the
 * localised position, mass, type (TYPE_STAR 1, DM 2, GAS 3, SINK 4),
 * link02 (gfind.c:51) and velocity; per halo the particle offsets
 * (nhalo + 1), the byte offsets of its records in pgalf_catalog_records
 * (nhalo + 1), its HaloQ, its Dump2BP origin (3 doubles) and its grid. */
int pgalf_catalog_fetch(const pgalf_catalog_t *cat, double *x, double *y, double *z,
                        double *mass, int *type, float *link02, double *vx, double *vy,
                        double *vz, long long *offsets, long long *record_offsets,
                        pgalf_haloq_t *haloq, double *origin, pgalf_grid_t *grids);
/* The last batch's records as read (valid until the next call). */
const void *pgalf_catalog_records(const pgalf_catalog_t *cat, long long *bytes);
void pgalf_catalog_close(pgalf_catalog_t *cat);

typedef struct pgalf_writer pgalf_writer_t;
/* Creates the output files (background may be NULL: no background file,
 * as -DNOBACKGROUND). */
pgalf_writer_t *pgalf_writer_open(const char *galfind_data, const char *catalog_list,
                                  const char *background);
/* write_data for one halo: its np particles' types and records
 * (concatenated in particle order, each of its type's size), the subhalo
 * of each particle (p2halo, -9999 for none) and the subhalo count.  Nothing
 * is written when mpeak <= 0. */
int pgalf_writer_halo(pgalf_writer_t *w, int np, const int *type, const void *records,
                      const int *p2halo, int mpeak);
int pgalf_writer_close(pgalf_writer_t *w);

#ifdef __cplusplus
}
#endif

#endif
