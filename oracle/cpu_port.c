/*
 * FP32 OpenMP restatement of IBIsCO's step loop for non-hybrid systems with
 * bonds and angles: the CPU baseline that bench.py times beside the GPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded by bench.py's cpu_baseline leg and by
 * tests/; the product never links or calls it.
 *
 * It follows the reference's algorithm and threading, not the oracle's FP64
 * arithmetic:
 *   - REAL*4 positions, velocities, forces, tables and pressure tensor, REAL*8
 *     energy sums (module_var.F90:14-24,52-56);
 *   - SHIFT.f90:20-48 with OMP SCHEDULE(STATIC,1);
 *   - LINKS.f90:5-25 head-insertion cells (serial), MAPS.f90:20-77 13-cell
 *     half stencil, NEW_NEIGHBOUR_WITHLIST.f90:5-127 with OMP SCHEDULE(STATIC,1)
 *     into LIST(MAXNAB, N) columns (the per-rebuild zeroing of LIST at :27 is
 *     left out: it does not change any result);
 *   - NONBONDED_FORCE.F90:35-105: OMP SCHEDULE(STATIC,1) with a whole-array
 *     REDUCTION(+:FXYZ) and scalar reductions of V_NB and the six PT terms;
 *   - BONDED_FORCE.F90:38-105 (BONDS.inc, ANGLES.inc): a parallel region with a
 *     private FXYZ_L per thread, OMP DO SCHEDULE(STATIC,1) and REDUCTION(+:FXYZ);
 *   - the bonded virial of NEW_FORCE.F90:73-80 and MOVE.F90:9-118 serially.
 * Set-up (FTABLE'd tables, SETLIS angles, CONNECTED_TO lists) is taken from an
 * FP64 oracle context of the same system.  Groups with fewer than 3 cells per
 * dimension use the all-pairs build, as the oracle does.
 */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ox_private.h"

typedef struct {
    int ndat;
    float bin;
    float *x, *V, *F;   /* ndat+2 rows */
} ftab_t;

typedef struct cp_ctx {
    int n, ntype, nupdate, halt_drift, ensemble;
    float box[3], boxinv[3], box2[3], dt, fac, temp_in;
    float rcut, rlist;
    float *s, *r, *v, *vt, *f, *fnb, *mass, *invmass;
    int *itype;
    int nnbt, nbt, nat;
    ftab_t *nt, *bt, *at;
    int *inbondt, *ibondt, *iangt;
    int *nbonds, *jbond;         /* 4 slots per particle */
    int *ang_off, *ang_j, *ang_k;
    int *conn_off, *conn;
    /* lists */
    int nc[3], ncell, allpairs;
    int *head, *lcl, *cell, *map, *list, *nneb;
    /* observables */
    double vnb, vbond, vang, ek;
    float pt[6];
    float temp;
    float mktemp;
} cp_ctx;

static void copy_tabs(const tab_t *src, int n, ftab_t **dst) {
    *dst = (ftab_t *)calloc((size_t)(n ? n : 1), sizeof(ftab_t));
    for (int t = 0; t < n; ++t) {
        ftab_t *d = &(*dst)[t];
        const int rows = src[t].ndat + 2;
        d->ndat = src[t].ndat;
        d->bin = (float)src[t].bin;
        d->x = (float *)malloc(sizeof(float) * rows);
        d->V = (float *)malloc(sizeof(float) * rows);
        d->F = (float *)malloc(sizeof(float) * rows);
        for (int k = 0; k < rows; ++k) {
            d->x[k] = (float)src[t].x[k];
            d->V[k] = (float)src[t].V[k];
            d->F[k] = (float)src[t].F[k];
        }
    }
}

static int *copy_ints(const int *src, size_t n) {
    int *d = (int *)malloc(sizeof(int) * (n ? n : 1));
    if (n) memcpy(d, src, sizeof(int) * n);
    return d;
}

void cp_destroy(cp_ctx *c);

cp_ctx *cp_create(const ox_ctx *o, char *err, int errlen) {
    if (o->hybrid || o->nvs || o->ntt || o->nopt || o->numatoms != o->natoms) {
        snprintf(err, (size_t)errlen, "the CPU port covers non-hybrid systems with bonds and angles only");
        return NULL;
    }
    cp_ctx *c = (cp_ctx *)calloc(1, sizeof(cp_ctx));
    const int n = o->natoms;
    c->n = n;
    c->ntype = o->ntype;
    c->nupdate = o->nupdate;
    c->halt_drift = o->halt_drift;
    c->ensemble = o->ensemble;
    c->dt = (float)o->dt;
    c->fac = (float)o->fac;
    c->temp_in = (float)o->temp_in;
    c->rcut = (float)o->g[0].rcut;
    c->rlist = (float)o->g[0].rlist;
    for (int k = 0; k < 3; ++k) {
        c->box[k] = (float)o->box[k];
        c->boxinv[k] = 1.0f / c->box[k];
        c->box2[k] = 0.5f * c->box[k];
    }
    c->s = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    c->r = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    c->v = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    c->vt = (float *)calloc(3 * (size_t)n, sizeof(float));
    c->f = (float *)calloc(3 * (size_t)n, sizeof(float));
    c->fnb = (float *)calloc(3 * (size_t)n, sizeof(float));
    for (size_t i = 0; i < 3 * (size_t)n; ++i) {
        c->s[i] = (float)o->s[i];
        c->v[i] = (float)o->v[i];
    }
    c->mass = (float *)malloc(sizeof(float) * o->ntype);
    c->invmass = (float *)malloc(sizeof(float) * o->ntype);
    for (int t = 0; t < o->ntype; ++t) {
        c->mass[t] = (float)o->mass[t];
        c->invmass[t] = 1.0f / c->mass[t];
    }
    c->itype = copy_ints(o->itype, n);
    c->nnbt = o->nnbt; c->nbt = o->nbt; c->nat = o->nat;
    copy_tabs(o->nt, o->nnbt, &c->nt);
    copy_tabs(o->bt, o->nbt, &c->bt);
    copy_tabs(o->at, o->nat, &c->at);
    const size_t nt2 = (size_t)o->ntype * o->ntype, nt3 = nt2 * o->ntype;
    c->inbondt = copy_ints(o->inbondt, nt2);
    c->ibondt = copy_ints(o->ibondt, nt2);
    c->iangt = copy_ints(o->iangt, nt3);
    c->nbonds = copy_ints(o->nbonds, n);
    c->jbond = copy_ints(o->jbond, 4 * (size_t)n);
    c->ang_off = copy_ints(o->ang.off, (size_t)n + 1);
    c->ang_j = copy_ints(o->ang.a, o->ang.off[n]);
    c->ang_k = copy_ints(o->ang.b, o->ang.off[n]);
    c->conn_off = copy_ints(o->conn_off, (size_t)n + 1);
    c->conn = copy_ints(o->conn, o->conn_off[n]);
    c->mktemp = (float)o->mktemp;
    c->temp = (float)o->temp;   /* main.F90:179-195 */
    /* ALLOCATEVAR2.f90:19-27 cell counts; MAPS.f90:20-77 stencil */
    for (int k = 0; k < 3; ++k) c->nc[k] = (int)(c->box[k] / c->rlist);
    c->allpairs = c->nc[0] < 3 || c->nc[1] < 3 || c->nc[2] < 3;
    c->ncell = c->allpairs ? 1 : c->nc[0] * c->nc[1] * c->nc[2];
    c->head = (int *)malloc(sizeof(int) * (size_t)c->ncell);
    c->lcl = (int *)malloc(sizeof(int) * (size_t)n);
    c->cell = (int *)malloc(sizeof(int) * (size_t)n);
    c->map = (int *)malloc(sizeof(int) * 13 * (size_t)c->ncell);
    c->list = (int *)malloc(sizeof(int) * (size_t)MAXNAB * n);
    c->nneb = (int *)calloc((size_t)n, sizeof(int));
    if (!c->list) {
        snprintf(err, (size_t)errlen, "no memory for LIST(%d, %d)", MAXNAB, n);
        cp_destroy(c);
        return NULL;
    }
    if (!c->allpairs) {
        static const int st[13][3] = {{1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {-1, 1, 0}, {1, 0, -1}, {1, 1, -1},
                                      {0, 1, -1}, {-1, 1, -1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}, {-1, 1, 1},
                                      {0, 0, 1}};
        const int *nc = c->nc;
        for (int ix = 0; ix < nc[0]; ++ix)
            for (int iy = 0; iy < nc[1]; ++iy)
                for (int iz = 0; iz < nc[2]; ++iz) {
                    const int base = ((ix * nc[1] + iy) * nc[2] + iz) * 13;
                    for (int q = 0; q < 13; ++q) {
                        const int x = (ix + st[q][0] + nc[0]) % nc[0];
                        const int y = (iy + st[q][1] + nc[1]) % nc[1];
                        const int z = (iz + st[q][2] + nc[2]) % nc[2];
                        c->map[base + q] = (x * nc[1] + y) * nc[2] + z;
                    }
                }
    }
    return c;
}

void cp_destroy(cp_ctx *c) {
    if (!c) return;
    ftab_t *tabs[3] = {c->nt, c->bt, c->at};
    const int ntab[3] = {c->nnbt, c->nbt, c->nat};
    for (int q = 0; q < 3; ++q) {
        for (int t = 0; tabs[q] && t < ntab[q]; ++t) { free(tabs[q][t].x); free(tabs[q][t].V); free(tabs[q][t].F); }
        free(tabs[q]);
    }
    free(c->s); free(c->r); free(c->v); free(c->vt); free(c->f); free(c->fnb); free(c->mass); free(c->invmass);
    free(c->itype); free(c->inbondt); free(c->ibondt); free(c->iangt); free(c->nbonds); free(c->jbond);
    free(c->ang_off); free(c->ang_j); free(c->ang_k); free(c->conn_off); free(c->conn);
    free(c->head); free(c->lcl); free(c->cell); free(c->map); free(c->list); free(c->nneb);
    free(c);
}

/* SHIFT.f90:20-48 */
static void shift(cp_ctx *c) {
    const int n = c->n;
#pragma omp parallel for schedule(static, 1)
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const float s = c->s[3 * i + k];
            float r = s;
            if (s > c->box2[k] || s < -c->box2[k]) {
                const float x = s / c->box2[k];
                const float sp = x - 2.0f * (float)(int)(x / 2.0f);
                r = (sp - 2.0f * (float)(int)sp) * c->box2[k];
            }
            c->r[3 * i + k] = r;
        }
}

static int connected(const cp_ctx *c, int i, int j) {
    for (int p = c->conn_off[i]; p < c->conn_off[i + 1]; ++p)
        if (c->conn[p] == j) return 1;
    return 0;
}

static float pair_r2(const cp_ctx *c, int i, int j) {
    float r2 = 0.0f;
    for (int k = 0; k < 3; ++k) {
        float d = c->r[3 * i + k] - c->r[3 * j + k];
        d = d - roundf(d * c->boxinv[k]) * c->box[k];   /* ANINT */
        r2 += d * d;
    }
    return r2;
}

/* LINKS.f90:5-25 then NEW_NEIGHBOUR_WITHLIST.f90:5-127 */
static int rebuild(cp_ctx *c) {
    const int n = c->n;
    const float rl2 = c->rlist * c->rlist;
    int overflow = 0;
    if (c->allpairs) {
#pragma omp parallel for schedule(static, 1) reduction(| : overflow)
        for (int a = 0; a < n; ++a) {
            int cnt = 0;
            int *col = c->list + (size_t)MAXNAB * a;
            for (int b = a + 1; b < n && !overflow; ++b) {
                if (connected(c, a, b)) continue;
                if (pair_r2(c, a, b) < rl2) {
                    col[cnt++] = b;
                    if (cnt == MAXNAB) overflow = 1;
                }
            }
            c->nneb[a] = cnt;
        }
        return overflow ? -1 : 0;
    }
    const int *nc = c->nc;
    for (int q = 0; q < c->ncell; ++q) c->head[q] = -1;
    for (int a = 0; a < n; ++a) {
        int ci[3];
        for (int k = 0; k < 3; ++k) {
            ci[k] = (int)((c->r[3 * a + k] * c->boxinv[k] + 0.5f) * nc[k]);
            if (ci[k] >= nc[k]) ci[k] = nc[k] - 1;   /* x = +L/2 exactly */
            if (ci[k] < 0) ci[k] = 0;
        }
        const int jc = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
        c->lcl[a] = c->head[jc];
        c->head[jc] = a;
        c->cell[a] = jc;
    }
#pragma omp parallel for schedule(static, 1) reduction(| : overflow)
    for (int a = 0; a < n; ++a) {
        int cnt = 0;
        int *col = c->list + (size_t)MAXNAB * a;
        for (int b = c->lcl[a]; b >= 0; b = c->lcl[b]) {
            if (connected(c, a, b)) continue;
            if (pair_r2(c, a, b) < rl2) {
                col[cnt++] = b;
                if (cnt == MAXNAB) { overflow = 1; break; }
            }
        }
        for (int q = 0; q < 13 && !overflow; ++q)
            for (int b = c->head[c->map[c->cell[a] * 13 + q]]; b >= 0; b = c->lcl[b]) {
                if (connected(c, a, b)) continue;
                if (pair_r2(c, a, b) < rl2) {
                    col[cnt++] = b;
                    if (cnt == MAXNAB) { overflow = 1; break; }
                }
            }
        c->nneb[a] = cnt;
    }
    return overflow ? -1 : 0;
}

/* NONBONDED_FORCE.F90:35-105 */
static void nonbonded(cp_ctx *c) {
    const int n = c->n;
    const float rc2 = c->rcut * c->rcut;
    float *f = c->f;
    double vnb = 0.0;
    float p11 = 0, p22 = 0, p33 = 0, p12 = 0, p13 = 0, p23 = 0;
#pragma omp parallel for schedule(static, 1) reduction(+ : vnb, p11, p22, p33, p12, p13, p23) \
    reduction(+ : f[0:3 * n])
    for (int a = 0; a < n; ++a) {
        const int ti = c->itype[a];
        const float ri[3] = {c->r[3 * a], c->r[3 * a + 1], c->r[3 * a + 2]};
        float fi[3] = {0, 0, 0};
        const int *col = c->list + (size_t)MAXNAB * a;
        for (int q = 0; q < c->nneb[a]; ++q) {
            const int j = col[q];
            const ftab_t *t = &c->nt[c->inbondt[ti * c->ntype + c->itype[j]]];
            float d[3];
            for (int k = 0; k < 3; ++k) {
                d[k] = ri[k] - c->r[3 * j + k];
                d[k] = d[k] - roundf(d[k] * c->boxinv[k]) * c->box[k];
            }
            const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(r2 < rc2)) continue;
            const float rij = sqrtf(r2);
            int ni = (int)(rij / t->bin);
            if (ni > t->ndat) ni = t->ndat;
            const float alpha = (rij - t->x[ni]) / t->bin;
            const float fij = t->F[ni] * (1.0f - alpha) + t->F[ni + 1] * alpha;
            const float vij = t->V[ni] * (1.0f - alpha) + t->V[ni + 1] * alpha;
            vnb += vij;
            for (int k = 0; k < 3; ++k) {
                fi[k] += fij * d[k];
                f[3 * j + k] -= fij * d[k];
            }
            p11 += fij * d[0] * d[0];
            p22 += fij * d[1] * d[1];
            p33 += fij * d[2] * d[2];
            p12 += fij * d[1] * d[0];
            p13 += fij * d[2] * d[0];
            p23 += fij * d[2] * d[1];
        }
        for (int k = 0; k < 3; ++k) f[3 * a + k] += fi[k];
    }
    c->vnb = vnb;
    c->pt[0] += p11; c->pt[1] += p22; c->pt[2] += p33; c->pt[3] += p12; c->pt[4] += p13; c->pt[5] += p23;
}

/* BONDED_FORCE.F90:38-105 with BONDS.inc:3-57 and ANGLES.inc:2-75 */
static void bonded(cp_ctx *c) {
    const int n = c->n, nt = c->ntype;
    const float R2D = (float)(180.0 / 3.14159265358);
    float *f = c->f;
    double vb = 0.0, va = 0.0;
#pragma omp parallel reduction(+ : vb, va) reduction(+ : f[0:3 * n])
    {
        float *fl = (float *)calloc(3 * (size_t)n, sizeof(float));
#pragma omp for schedule(static, 1)
        for (int i = 0; i < n; ++i) {
            const float *ri = c->s + 3 * i;
            const int ti = c->itype[i];
            float fi[3] = {0, 0, 0};
            for (int b = 0; b < c->nbonds[i]; ++b) {
                const int j = c->jbond[4 * i + b];
                if (j <= i) continue;
                const ftab_t *t = &c->bt[c->ibondt[ti * nt + c->itype[j]]];
                float d[3];
                for (int k = 0; k < 3; ++k) d[k] = ri[k] - c->s[3 * j + k];
                const float rij = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                int ni = (int)((rij - t->x[0]) / t->bin);
                if (ni > t->ndat) ni = t->ndat;
                else if (ni < 0) ni = 0;
                const float alpha = (rij - t->x[ni]) / t->bin;
                const float fij = t->F[ni] * (1.0f - alpha) + alpha * t->F[ni + 1];
                vb += t->V[ni] * (1.0f - alpha) + alpha * t->V[ni + 1];
                for (int k = 0; k < 3; ++k) {
                    fi[k] += fij * d[k];
                    fl[3 * j + k] -= fij * d[k];
                }
            }
            for (int p = c->ang_off[i]; p < c->ang_off[i + 1]; ++p) {
                const int j = c->ang_j[p], kk = c->ang_k[p];
                const ftab_t *t = &c->at[c->iangt[(ti * nt + c->itype[j]) * nt + c->itype[kk]]];
                float ij[3], kj[3], M[3], N[3], S[3];
                for (int k = 0; k < 3; ++k) {
                    ij[k] = ri[k] - c->s[3 * j + k];
                    kj[k] = c->s[3 * kk + k] - c->s[3 * j + k];
                }
                const float rij = sqrtf(ij[0] * ij[0] + ij[1] * ij[1] + ij[2] * ij[2]);
                const float rkj = sqrtf(kj[0] * kj[0] + kj[1] * kj[1] + kj[2] * kj[2]);
                float ct = (ij[0] * kj[0] + ij[1] * kj[1] + ij[2] * kj[2]) / rij / rkj;
                if (ct > 1.0f) ct = 1.0f;
                if (ct < -1.0f) ct = -1.0f;
                const float theta = acosf(ct) * R2D;
                M[0] = -(ij[1] * kj[2] - ij[2] * kj[1]);
                M[1] = -(-ij[0] * kj[2] + ij[2] * kj[0]);
                M[2] = -(ij[0] * kj[1] - ij[1] * kj[0]);
                const float rm = sqrtf(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]);
                for (int k = 0; k < 3; ++k) M[k] /= rm;
                N[0] = (M[1] * ij[2] - M[2] * ij[1]) / rij;
                N[1] = (-M[0] * ij[2] + M[2] * ij[0]) / rij;
                N[2] = (M[0] * ij[1] - M[1] * ij[0]) / rij;
                S[0] = -(M[1] * kj[2] - M[2] * kj[1]) / rkj;
                S[1] = -(-M[0] * kj[2] + M[2] * kj[0]) / rkj;
                S[2] = -(M[0] * kj[1] - M[1] * kj[0]) / rkj;
                int ni = (int)(theta / t->bin);
                if (!(ni >= 0 && ni <= t->ndat)) ni = ni < 0 ? 0 : t->ndat;   /* NaN geometry of a blown-up run */
                float fa, ve;
                if (theta == 180.0f) { fa = t->F[ni]; ve = t->V[ni]; }
                else {
                    const float alpha = (theta - t->x[ni]) / t->bin;
                    fa = t->F[ni] * (1.0f - alpha) + alpha * t->F[ni + 1];
                    ve = t->V[ni] * (1.0f - alpha) + alpha * t->V[ni + 1];
                }
                va += ve;
                for (int k = 0; k < 3; ++k) {
                    const float f1 = fa * N[k] / rij, f3 = fa * S[k] / rkj;
                    fi[k] += f1;
                    fl[3 * kk + k] += f3;
                    fl[3 * j + k] -= f3 + f1;
                }
            }
            for (int k = 0; k < 3; ++k) fl[3 * i + k] += fi[k];
        }
        for (size_t q = 0; q < 3 * (size_t)n; ++q) f[q] += fl[q];
        free(fl);
    }
    c->vbond = vb;
    c->vang = va;
}

/* NEW_FORCE.F90:7-83 for one group of atoms */
static void new_force(cp_ctx *c) {
    const size_t m = 3 * (size_t)c->n;
    memset(c->f, 0, sizeof(float) * m);
    for (int k = 0; k < 6; ++k) c->pt[k] = 0.0f;
    nonbonded(c);
    memcpy(c->fnb, c->f, sizeof(float) * m);
    bonded(c);
    float w[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < c->n; ++i) {
        const float *s = c->s + 3 * i;
        float d[3];
        for (int k = 0; k < 3; ++k) d[k] = c->f[3 * i + k] - c->fnb[3 * i + k];
        w[0] += d[0] * s[0]; w[1] += d[1] * s[1]; w[2] += d[2] * s[2];
        w[3] += d[0] * s[1]; w[4] += d[0] * s[2]; w[5] += d[1] * s[2];
    }
    for (int k = 0; k < 6; ++k) c->pt[k] += w[k];
}

/* MOVE.F90:9-118 (one population) */
static void move(cp_ctx *c) {
    const float lam = (c->ensemble == 1) ? sqrtf(1.0f + c->fac * ((c->temp_in / c->temp) - 1.0f)) : 1.0f;
    float pk[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < c->n; ++i) {
        const float m = c->mass[c->itype[i]], im = c->invmass[c->itype[i]];
        float vt[3];
        for (int k = 0; k < 3; ++k) {
            const float v0 = c->v[3 * i + k];
            const float v1 = (v0 + c->dt * c->f[3 * i + k] * im) * lam;
            c->v[3 * i + k] = v1;
            c->s[3 * i + k] += c->dt * v1;
            vt[k] = 0.5f * (v1 + v0);
            c->vt[3 * i + k] = vt[k];
        }
        pk[0] += m * vt[0] * vt[0]; pk[1] += m * vt[1] * vt[1]; pk[2] += m * vt[2] * vt[2];
        pk[3] += m * vt[0] * vt[1]; pk[4] += m * vt[0] * vt[2]; pk[5] += m * vt[1] * vt[2];
    }
    c->ek = 0.5 * ((double)pk[0] + pk[1] + pk[2]);
    c->temp = (float)(c->ek * c->mktemp);
    const float vol = c->box[0] * c->box[1] * c->box[2];
    for (int k = 0; k < 6; ++k) c->pt[k] = (c->pt[k] + pk[k]) / vol;
}

/* MOMENTUM.f90:5-32 */
static void momentum(cp_ctx *c) {
    float p[3] = {0, 0, 0};
    for (int i = 0; i < c->n; ++i)
        for (int k = 0; k < 3; ++k) p[k] += c->v[3 * i + k] * c->mass[c->itype[i]];
    for (int k = 0; k < 3; ++k) p[k] /= (float)c->n;
    for (int i = 0; i < c->n; ++i)
        for (int k = 0; k < 3; ++k) c->v[3 * i + k] -= p[k] * c->invmass[c->itype[i]];
}

static double now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* The first list build (main.F90:234). */
int cp_init_lists(cp_ctx *c) {
    shift(c);
    return rebuild(c);
}

/* nsteps of NEW_LOOP.F90:19-156 after step0; *seconds gets the loop's wall time. */
int cp_run(cp_ctx *c, int nsteps, int step0, double *seconds) {
    const double t0 = now();
    for (int step = step0 + 1; step <= step0 + nsteps; ++step) {
        shift(c);
        if (step % c->nupdate == 0 && rebuild(c)) return -1;
        new_force(c);
        move(c);
        if (step % c->halt_drift == 0) momentum(c);
    }
    if (seconds) *seconds = now() - t0;
    return 0;
}

int cp_threads(void) { return omp_get_max_threads(); }

void cp_get_forces(const cp_ctx *c, double *f) {
    for (size_t q = 0; q < 3 * (size_t)c->n; ++q) f[q] = c->f[q];
}

/* obs as ox_get_obs for the fields the port computes: V_NB(1), V_BOND(1),
 * V_ANGLE(1), PT11..PT23, EK(1), TEMP; the rest 0. */
void cp_get_obs(const cp_ctx *c, double *o) {
    memset(o, 0, sizeof(double) * 32);
    o[0] = c->vnb;
    o[3] = c->vbond;
    o[6] = c->vang;
    for (int k = 0; k < 6; ++k) o[15 + k] = c->pt[k];
    o[21] = c->ek;
    o[23] = c->temp;
}

void cp_get_state(const cp_ctx *c, double *s, double *v) {
    for (size_t q = 0; q < 3 * (size_t)c->n; ++q) {
        if (s) s[q] = c->s[q];
        if (v) v[q] = c->v[q];
    }
}
