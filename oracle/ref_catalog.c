/*
 * ref_catalog.c -- thin C driver around the REFERENCE's own catalogue code
 * (NewGalFinder/gfind.c FREAD and write_data, subhaloden.mod6.c Dump2BP and
 * mklocalize, built in place from /root/reference by oracle/Makefile).
 *
 * TEST INFRASTRUCTURE ONLY: pins the record layouts, the loader
 * (read + localisation) and the writer of galaxyfinder_amd's catalogue
 * code (tests/golden/make_golden.py catalog, tests/test_catalog.py).  The
 * loops over halos mirror gfind.c's master loop (read a halo, hand it on
 * when it has at least 30 particles, write what comes back).
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "Memory.h"
#include "header.h"
#include "ramses.h"
#include "tree.h"
#include "defs.h"

void FREAD(FoFTPtlStruct *, HaloQ *, FILE *, float);
void write_data(FoFTPtlStruct *, lint *, int);
double Dump2BP(FoFTPtlStruct *, SimpleBasicParticleType *, lint);
void mklocalize(SimpleBasicParticleType *, lint, float *, float *, float *, float *, float *,
                float *);
int ref_init(long long bytes);
extern FILE *wrp, *wbp, *wlist;
extern FoFTPtlStruct *rbuffer;
extern size_t nbuffer;
extern float size, omep; /* gfind.c globals (defs.h renames them) */
extern int mpeak;

/* sizeof / offsetof of the catalogue records under the reference's flags */
void ref_catalog_layouts(long long *o) {
    o[0] = sizeof(HaloQ);
    o[1] = sizeof(DmType);
    o[2] = sizeof(GasType);
    o[3] = sizeof(SinkType);
    o[4] = sizeof(StarType);
    o[5] = sizeof(HaloInfo);
    o[6] = sizeof(SubInfo);
    o[7] = offsetof(HaloQ, x);
    o[8] = offsetof(HaloQ, mass);
    o[9] = offsetof(PmType, mass);
    o[10] = offsetof(PmType, id);
    o[11] = offsetof(PmType, tp);
    o[12] = offsetof(PmType, partp);
    o[13] = offsetof(GasType, vx);
    o[14] = offsetof(GasType, temp);
    o[15] = offsetof(GasType, mass);
    o[16] = offsetof(GasType, potent);
    o[17] = offsetof(SinkType, mass);
    o[18] = offsetof(SinkType, id);
    o[19] = offsetof(HaloInfo, totm);
    o[20] = offsetof(SubInfo, totm);
    o[21] = offsetof(HaloQ, vx);
}

static void ref_buffers(void) {
    if (!rbuffer) {
        nbuffer = 1;
        rbuffer = (FoFTPtlStruct *)Malloc(sizeof(FoFTPtlStruct) * nbuffer, PPTR(rbuffer));
    }
}

/* Every halo of the catalogue (up to max_halos) read by FREAD, packed by
 * Dump2BP and unwrapped by mklocalize, as subhalo_den does
 * (subhaloden.mod6.c:2874,2883), with the box size and omep of the header.
 * Outputs concatenated (at most cap particles): localised x, y, z, mass,
 * type, link02; per halo its offset, the Dump2BP origin, and np. */
long long ref_read_catalog(const char *hcat, const char *members, int max_halos, long long cap,
                           double *x, double *y, double *z, double *m, int *type, float *link02,
                           long long *off, double *origin) {
    ref_init(4LL << 30);
    ref_buffers();
    FILE *hf = fopen(hcat, "rb"), *mf = fopen(members, "rb");
    if (!hf || !mf) return -1;
    float hdr[7];
    if (fread(hdr, sizeof(float), 7, hf) != 7) return -1;
    size = hdr[0];
    omep = hdr[2];
    HaloQ q;
    long long nh = 0, tot = 0;
    off[0] = 0;
    while (nh < max_halos && fread(&q, sizeof(HaloQ), 1, hf) == 1) {
        FoFTPtlStruct *p = (FoFTPtlStruct *)calloc(q.np, sizeof(FoFTPtlStruct));
        SimpleBasicParticleType *bp =
            (SimpleBasicParticleType *)calloc(q.np, sizeof(SimpleBasicParticleType));
        FREAD(p, &q, mf, omep);
        double ox = 1e30, oy = 1e30, oz = 1e30;
        for (size_t i = 0; i < q.np; i++) {
            ox = p[i].x < ox ? p[i].x : ox;
            oy = p[i].y < oy ? p[i].y : oy;
            oz = p[i].z < oz ? p[i].z : oz;
        }
        Dump2BP(p, bp, (lint)q.np);
        float xi, yi, zi, xa, ya, za;
        mklocalize(bp, (lint)q.np, &xi, &yi, &zi, &xa, &ya, &za);
        for (size_t i = 0; i < q.np && tot < cap; i++, tot++) {
            x[tot] = bp[i].x;
            y[tot] = bp[i].y;
            z[tot] = bp[i].z;
            m[tot] = bp[i].mass;
            type[tot] = (int)bp[i].type;
            link02[tot] = p[i].link02;
        }
        origin[3 * nh] = ox;
        origin[3 * nh + 1] = oy;
        origin[3 * nh + 2] = oz;
        nh++;
        off[nh] = tot;
        free(p);
        free(bp);
    }
    fclose(hf);
    fclose(mf);
    return nh;
}

/* gfind.c's master loop on a catalogue: every halo of at least 30
 * particles read by FREAD and written by write_data with the given
 * membership (p2halo over the whole catalogue, concatenated) and subhalo
 * count per halo. */
long long ref_write_catalog(const char *hcat, const char *members, const int *p2halo,
                            const int *mpeaks, const char *galfind, const char *catlist,
                            const char *background) {
    ref_init(4LL << 30);
    ref_buffers();
    FILE *hf = fopen(hcat, "rb"), *mf = fopen(members, "rb");
    if (!hf || !mf) return -1;
    float hdr[7];
    if (fread(hdr, sizeof(float), 7, hf) != 7) return -1;
    size = hdr[0];
    omep = hdr[2];
    wrp = fopen(galfind, "wb");
    wlist = fopen(catlist, "wb");
    wbp = fopen(background, "wb");
    HaloQ q;
    long long nh = 0, tot = 0;
    while (fread(&q, sizeof(HaloQ), 1, hf) == 1) {
        FoFTPtlStruct *p = (FoFTPtlStruct *)calloc(q.np, sizeof(FoFTPtlStruct));
        FREAD(p, &q, mf, omep);
        if (q.np >= 30) {
            lint *p2 = (lint *)malloc(sizeof(lint) * q.np);
            for (size_t i = 0; i < q.np; i++) p2[i] = p2halo[tot + i];
            mpeak = mpeaks[nh];
            write_data(p, p2, (int)q.np);
            free(p2);
        }
        tot += (long long)q.np;
        nh++;
        free(p);
    }
    fclose(wrp);
    fclose(wlist);
    fclose(wbp);
    fclose(hf);
    fclose(mf);
    return nh;
}
