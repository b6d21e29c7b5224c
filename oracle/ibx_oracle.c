/*
 * FP64 CPU restatement of IBIsCO's per-step force and integration path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle for the HIP path in ibisco_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function cites the reference routine it follows.  Deliberate deviations
 * from the reference (all documented at the function):
 *   - double precision throughout (the reference is REAL*4);
 *   - O(N) exclusion sets instead of the NITEMS^2 IS_CONNECTED matrix
 *     (BUILD_CONNECTIVITY.f90:23), same resulting CONNECTED_TO lists;
 *   - cell index clamped at x = +L/2 (LINKS.f90:16-18 overflows one row);
 *   - groups with fewer than 3 cells per dimension use the all-pairs build
 *     (MAPS.f90:39-51 would list the same neighbour cell twice);
 *   - centre-of-mass virtual sites zero all three accumulators
 *     (VIRTUAL_DEF.f90:31 zeroes only x).
 */
#include "ibx_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ox_private.h"

static const double PI = 3.14159265358;   /* module_var.F90:16 */

static void iv_push(ivec *v, int x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 16;
        v->data = (int *)realloc(v->data, sizeof(int) * (size_t)v->cap);
    }
    v->data[v->n++] = x;
}

static int ti2(const ox_ctx *c, int a, int b) { return a * c->ntype + b; }
static int ti3(const ox_ctx *c, int a, int b, int d) { return (a * c->ntype + b) * c->ntype + d; }
static int ti4(const ox_ctx *c, int a, int b, int d, int e) {
    return ((a * c->ntype + b) * c->ntype + d) * c->ntype + e;
}

/* ANINT: round half away from zero */
static double anint(double x) { return round(x); }

/* RDINTERACT.f90:450-537 (table load with the shift when row 0 is not at 0),
 * restated for angle/torsion/nonbonded/OOP tables; bonds load rows as they are
 * (RDINTERACT.f90:205-220). */
static void load_table(tab_t *t, int nrow, const double *x, const double *v, int shift, int oop) {
    int cap = nrow + 4;
    t->x = (double *)calloc((size_t)cap, sizeof(double));
    t->V = (double *)calloc((size_t)cap, sizeof(double));
    t->F = (double *)calloc((size_t)cap, sizeof(double));
    double bin = x[1] - x[0];
    if (oop && x[1] < x[0]) bin = x[0] - x[1];
    int k;
    if (shift && x[0] != 0.0) {
        t->x[0] = 0.0; t->V[0] = v[0];
        t->x[1] = bin; t->V[1] = v[0];
        t->x[2] = x[1]; t->V[2] = v[1];
        for (k = 3; k - 1 < nrow; ++k) { t->x[k] = x[k - 1]; t->V[k] = v[k - 1]; }
    } else {
        for (k = 0; k < nrow; ++k) { t->x[k] = x[k]; t->V[k] = v[k]; }
    }
    t->ndat = k - 1;
    t->bin = bin;
}

/* FTABLE.f90:9-91.  Central differences; nonbonded and bond tables store F/r. */
static void ftable(ox_ctx *c) {
    const double D2R = PI / 180.0;
    for (int i = 0; i < c->nbt; ++i) {
        tab_t *t = &c->bt[i];
        for (int j = 1; j <= t->ndat - 1; ++j)
            t->F[j] = -(t->V[j + 1] - t->V[j - 1]) / t->bin / t->x[j] / 2.0;
        t->F[0] = t->F[1];
        /* FTABLE.f90:31 writes BOND_FORCE(I, NDATB(:)) with the whole array as a
         * vector subscript; the intended effect (last row = previous row) is kept. */
        t->F[t->ndat] = t->F[t->ndat - 1];
    }
    for (int i = 0; i < c->nat; ++i) {
        tab_t *t = &c->at[i];
        for (int j = 1; j <= t->ndat - 1; ++j)
            t->F[j] = -(t->V[j + 1] - t->V[j - 1]) / t->bin / D2R / 2.0;
        t->F[0] = t->F[1];
        t->F[t->ndat] = t->F[t->ndat - 1];
    }
    for (int i = 0; i < c->ntt; ++i) {
        tab_t *t = &c->tt[i];
        for (int j = 1; j <= t->ndat - 1; ++j)
            t->F[j] = -(t->V[j + 1] - t->V[j - 1]) / t->bin / D2R / 2.0;
        t->F[0] = t->F[1];
        t->F[t->ndat] = t->F[0];
    }
    for (int i = 0; i < c->nnbt; ++i) {
        tab_t *t = &c->nt[i];
        /* V[ndat+1] is the zero row (FTABLE.f90:61-62 reads past the data) */
        for (int j = 1; j <= t->ndat; ++j)
            t->F[j] = -(t->V[j + 1] - t->V[j - 1]) / t->bin / t->x[j] / 2.0;
        t->F[0] = t->F[1];
    }
    for (int i = 0; i < c->nopt; ++i) {
        tab_t *t = &c->ot[i];
        for (int j = 1; j <= t->ndat - 1; ++j)
            t->F[j] = -(t->V[j + 1] - t->V[j - 1]) / t->bin / D2R / 2.0;
        t->F[0] = t->F[1];
        t->F[t->ndat] = t->F[0];
    }
}

/* label hash of RDINTERACT.f90:621-701: 1 all atoms, 2 all beads, 3 mixed */
static int type_class(const ox_ctx *c, const int *types, int k) {
    int h = 0;
    for (int i = 0; i < k; ++i) h += c->tlabel[types[i]];
    if (h == k) return 1;
    if (h == 2 * k) return 2;
    return 3;
}

static int setup_tables(ox_ctx *c, const ox_tables *src, int k, tab_t **dst, int *count, int **map,
                        int **labels, int shift, int oop, int symmetric) {
    int nt = c->ntype;
    size_t msz = 1;
    for (int i = 0; i < k; ++i) msz *= (size_t)nt;
    *map = (int *)malloc(sizeof(int) * msz);
    for (size_t i = 0; i < msz; ++i) (*map)[i] = NONE;
    *count = src->ntab;
    *dst = (tab_t *)calloc((size_t)(src->ntab > 0 ? src->ntab : 1), sizeof(tab_t));
    *labels = (int *)calloc((size_t)(src->ntab > 0 ? src->ntab : 1), sizeof(int));
    for (int t = 0; t < src->ntab; ++t) {
        const int *ty = src->types + t * k;
        size_t fwd = 0, bwd = 0;
        for (int i = 0; i < k; ++i) fwd = fwd * nt + ty[i];
        for (int i = k - 1; i >= 0; --i) bwd = bwd * nt + ty[i];
        if ((*map)[fwd] != NONE) {
            snprintf(c->err, sizeof c->err, "duplicate table entry for types (kind %d)", k);
            return -1;
        }
        (*map)[fwd] = t;
        if (symmetric) (*map)[bwd] = t;
        load_table(&(*dst)[t], src->nrow[t], src->x + src->off[t], src->v + src->off[t], shift, oop);
        (*labels)[t] = type_class(c, ty, k);
    }
    return 0;
}

static void terms_build(terms_t *T, ivec *own, ivec *a, ivec *b, ivec *cc, int nitems) {
    T->n = own->n;
    T->off = (int *)calloc((size_t)nitems + 1, sizeof(int));
    T->a = (int *)malloc(sizeof(int) * (size_t)(own->n + 1));
    T->b = (int *)malloc(sizeof(int) * (size_t)(own->n + 1));
    T->c = (int *)malloc(sizeof(int) * (size_t)(own->n + 1));
    for (int i = 0; i < own->n; ++i) T->off[own->data[i] + 1]++;
    for (int i = 0; i < nitems; ++i) T->off[i + 1] += T->off[i];
    int *fill = (int *)calloc((size_t)nitems, sizeof(int));
    for (int i = 0; i < own->n; ++i) {   /* stable: keeps enumeration order per owner */
        int o = own->data[i];
        int p = T->off[o] + fill[o]++;
        T->a[p] = a->data[i];
        T->b[p] = b ? b->data[i] : -1;
        T->c[p] = cc ? cc->data[i] : -1;
    }
    free(fill);
}

/* SETLIS.f90:26-419: angles, torsions and out-of-plane terms derived from the
 * bonds.  Angle i-j-k is owned by the end atom i with k>i, torsion i-j-k-l by i
 * with l>i.  Terms with no table only feed the exclusions. */
static int setlis(ox_ctx *c, const ox_desc *d) {
    ivec ao = {0}, aj = {0}, ak = {0}, no = {0}, nj = {0}, nk = {0};
    ivec to = {0}, tj = {0}, tk = {0}, tl = {0}, fo = {0}, fj = {0}, fk = {0}, fl = {0};
    ivec oo = {0}, oj = {0}, ok = {0}, ol = {0};
    if (c->nbt > 0) {
        for (int I = 0; I < c->natoms; ++I) {
            int IT = c->itype[I];
            const int *jb = c->jbond + 4 * I;
            if (c->nopt > 0 && c->nbonds[I] == 3) {
                static const int perm[6][3] = {{0, 1, 2}, {2, 0, 1}, {1, 0, 2}, {0, 2, 1}, {1, 2, 0}, {2, 1, 0}};
                int found = 0;
                for (int p = 0; p < 6; ++p) {
                    int Z = jb[perm[p][0]], P = jb[perm[p][1]], O = jb[perm[p][2]];
                    if (c->ioopt[ti4(c, IT, c->itype[Z], c->itype[P], c->itype[O])] != NONE && found == 0) {
                        found = 1;
                        iv_push(&oo, I); iv_push(&oj, Z); iv_push(&ok, P); iv_push(&ol, O);
                    }
                }
            }
            for (int j1 = 0; j1 < c->nbonds[I]; ++j1) {
                int J = jb[j1], JT = c->itype[J];
                if (c->ibondt[ti2(c, IT, JT)] == NONE) {
                    snprintf(c->err, sizeof c->err, "no bond parameters for atoms %d-%d", I + 1, J + 1);
                    return -1;
                }
                if (c->nat <= 0) continue;
                for (int k1 = 0; k1 < c->nbonds[J]; ++k1) {
                    int K = c->jbond[4 * J + k1], KT = c->itype[K];
                    if (K == I) continue;
                    if (K > I) {
                        if (c->iangt[ti3(c, IT, JT, KT)] == NONE) {
                            iv_push(&no, I); iv_push(&nj, J); iv_push(&nk, K);
                        } else {
                            iv_push(&ao, I); iv_push(&aj, J); iv_push(&ak, K);
                        }
                    }
                    if (c->ntt <= 0) continue;
                    for (int l1 = 0; l1 < c->nbonds[K]; ++l1) {
                        int L = c->jbond[4 * K + l1], LT = c->itype[L];
                        if (L == J) continue;
                        if (L > I) {
                            if (c->itort[ti4(c, IT, JT, KT, LT)] == NONE) {
                                if (c->nopt > 0 && c->ioopt[ti4(c, IT, JT, KT, LT)] != NONE) {
                                    iv_push(&oo, I); iv_push(&oj, J); iv_push(&ok, K); iv_push(&ol, L);
                                } else {
                                    iv_push(&fo, I); iv_push(&fj, J); iv_push(&fk, K); iv_push(&fl, L);
                                }
                            } else {
                                iv_push(&to, I); iv_push(&tj, J); iv_push(&tk, K); iv_push(&tl, L);
                            }
                        }
                    }
                }
            }
        }
    }
    /* RDVIRTANGLES.f90:36-50 appends (j,k) to owner i after SETLIS. */
    for (int q = 0; q < d->nvirtangles; ++q) {
        iv_push(&ao, d->virtangles[3 * q]);
        iv_push(&aj, d->virtangles[3 * q + 1]);
        iv_push(&ak, d->virtangles[3 * q + 2]);
    }
    terms_build(&c->ang, &ao, &aj, &ak, NULL, c->nitems);
    terms_build(&c->noang, &no, &nj, &nk, NULL, c->nitems);
    terms_build(&c->tor, &to, &tj, &tk, &tl, c->nitems);
    terms_build(&c->ftor, &fo, &fj, &fk, &fl, c->nitems);
    terms_build(&c->oop, &oo, &oj, &ok, &ol, c->nitems);
    ivec *all[] = {&ao, &aj, &ak, &no, &nj, &nk, &to, &tj, &tk, &tl, &fo, &fj, &fk, &fl, &oo, &oj, &ok, &ol};
    for (size_t q = 0; q < sizeof all / sizeof all[0]; ++q) free(all[q]->data);
    return 0;
}

/* BUILD_CONNECTIVITY.f90:14-116 restated in O(N): CONNECTED_TO(I) lists, in
 * ascending order, every atom J sharing a bond, angle (or, with exclusion 5,
 * torsion/OOP) term with I, the term being owned by I or by J.  Then
 * RDVIRTBONDS.f90:42-64 appends the virtbonds pairs in file order. */
static int build_connectivity(ox_ctx *c, const ox_desc *d) {
    int n = c->nitems;
    ivec *base = (ivec *)calloc((size_t)n, sizeof(ivec));
    ivec *p14 = (ivec *)calloc((size_t)n, sizeof(ivec));
#define LINK(V, a, b) do { iv_push(&(V)[(a)], (b)); iv_push(&(V)[(b)], (a)); } while (0)
    for (int I = 0; I < c->natoms; ++I) {
        for (int b = 0; b < c->nbonds[I]; ++b) LINK(base, I, c->jbond[4 * I + b]);
        for (int p = c->ang.off[I]; p < c->ang.off[I + 1]; ++p) {
            LINK(base, I, c->ang.a[p]); LINK(base, I, c->ang.b[p]);
        }
        for (int p = c->noang.off[I]; p < c->noang.off[I + 1]; ++p) {
            LINK(base, I, c->noang.a[p]); LINK(base, I, c->noang.b[p]);
        }
        terms_t *T4[3] = {&c->tor, &c->ftor, &c->oop};
        for (int q = 0; q < 3; ++q)
            for (int p = T4[q]->off[I]; p < T4[q]->off[I + 1]; ++p) {
                LINK(p14, I, T4[q]->a[p]); LINK(p14, I, T4[q]->b[p]); LINK(p14, I, T4[q]->c[p]);
            }
    }
#undef LINK
    ivec *conn = (ivec *)calloc((size_t)n, sizeof(ivec));
    char *mark = (char *)calloc((size_t)n, 1);
    for (int I = 0; I < c->natoms; ++I) {
        int excl = (c->type_label[I] == 2) ? c->g[1].exclusion : c->g[0].exclusion;
        ivec tmp = {0};
        for (int q = 0; q < base[I].n; ++q) iv_push(&tmp, base[I].data[q]);
        if (excl == 5)
            for (int q = 0; q < p14[I].n; ++q) iv_push(&tmp, p14[I].data[q]);
        for (int q = 0; q < tmp.n; ++q) {
            int J = tmp.data[q];
            if (J < c->natoms) mark[J] = 1;
        }
        /* ascending scan restricted to the marked atoms */
        int cnt = 0;
        for (int q = 0; q < tmp.n; ++q) {
            int J = tmp.data[q];
            if (J < c->natoms && mark[J] == 1) { mark[J] = 2; cnt++; }
        }
        int *sorted = (int *)malloc(sizeof(int) * (size_t)(cnt + 1));
        int m = 0;
        for (int q = 0; q < tmp.n; ++q) {
            int J = tmp.data[q];
            if (J < c->natoms && mark[J] == 2) { sorted[m++] = J; mark[J] = 0; }
        }
        /* insertion sort (lists are short) */
        for (int a = 1; a < m; ++a) {
            int x = sorted[a], b = a - 1;
            while (b >= 0 && sorted[b] > x) { sorted[b + 1] = sorted[b]; --b; }
            sorted[b + 1] = x;
        }
        for (int a = 0; a < m; ++a) iv_push(&conn[I], sorted[a]);
        free(sorted);
        free(tmp.data);
    }
    for (int q = 0; q < d->nvirtbonds; ++q) {
        int I = d->virtbonds[2 * q], J = d->virtbonds[2 * q + 1];
        iv_push(&conn[I], J);
        iv_push(&conn[J], I);
    }
    c->conn_off = (int *)calloc((size_t)n + 1, sizeof(int));
    for (int I = 0; I < n; ++I) c->conn_off[I + 1] = c->conn_off[I] + conn[I].n;
    c->conn = (int *)malloc(sizeof(int) * (size_t)(c->conn_off[n] + 1));
    for (int I = 0; I < n; ++I) {
        if (conn[I].n) memcpy(c->conn + c->conn_off[I], conn[I].data, sizeof(int) * (size_t)conn[I].n);
        free(conn[I].data); free(base[I].data); free(p14[I].data);
    }
    free(conn); free(base); free(p14); free(mark);
    return 0;
}

void ox_destroy(ox_ctx *c);

ox_ctx *ox_create(const ox_desc *d, char *err, int errlen) {
    ox_ctx *c = (ox_ctx *)calloc(1, sizeof(ox_ctx));
    c->respa = 1;
    c->natoms = d->natoms; c->nvs = d->nvs; c->nitems = d->natoms + d->nvs; c->ntype = d->ntype;
    c->hybrid = d->nvs > 0;
    int n = c->nitems;
    c->mass = (double *)malloc(sizeof(double) * (size_t)c->ntype);
    c->invmass = (double *)malloc(sizeof(double) * (size_t)c->ntype);
    c->tlabel = (int *)malloc(sizeof(int) * (size_t)c->ntype);
    for (int t = 0; t < c->ntype; ++t) {
        c->mass[t] = d->mass[t];
        c->invmass[t] = 1.0 / d->mass[t];
        c->tlabel[t] = d->label[t];
    }
    c->itype = (int *)calloc((size_t)n, sizeof(int));
    c->type_label = (int *)calloc((size_t)n, sizeof(int));
    c->nbonds = (int *)calloc((size_t)n, sizeof(int));
    c->jbond = (int *)calloc((size_t)n * 4, sizeof(int));
    c->s = (double *)calloc((size_t)n * 3, sizeof(double));
    c->r = (double *)calloc((size_t)n * 3, sizeof(double));
    c->f = (double *)calloc((size_t)n * 3, sizeof(double));
    c->v = (double *)calloc((size_t)c->natoms * 3, sizeof(double));
    c->vt = (double *)calloc((size_t)c->natoms * 3, sizeof(double));
    c->fnb = (double *)calloc((size_t)c->natoms * 3, sizeof(double));
    for (int i = 0; i < c->natoms; ++i) {
        c->itype[i] = d->itype[i];
        /* RDCOOR.f90:84,110-123: labels come from the type only in hybrid mode */
        c->type_label[i] = c->hybrid ? d->label[d->itype[i]] : 1;
        c->nbonds[i] = d->nbonds[i];
        for (int k = 0; k < 4; ++k) c->jbond[4 * i + k] = d->jbond[4 * i + k];
    }
    memcpy(c->s, d->s, sizeof(double) * 3 * (size_t)c->natoms);
    memcpy(c->v, d->v, sizeof(double) * 3 * (size_t)c->natoms);
    for (int k = 0; k < 3; ++k) {
        c->box[k] = d->box[k];
        c->boxinv[k] = 1.0 / d->box[k];
        c->box2[k] = d->box[k] / 2.0;
    }
    c->ensemble = d->ensemble; c->dt = d->dt; c->temp_in = d->temp_in; c->fac = d->fac;
    c->pressure = d->pressure; c->pfac = d->pfac; c->iavp = 0;
    c->nupdate = d->nupdate; c->halt_drift = d->halt_drift; c->nmts = d->nmts;
    c->mts_quirk = d->mts_quirk;

    if (setup_tables(c, &d->bond, 2, &c->bt, &c->nbt, &c->ibondt, &c->bond_lab, 0, 0, 1) ||
        setup_tables(c, &d->angle, 3, &c->at, &c->nat, &c->iangt, &c->angle_lab, 1, 0, 1) ||
        setup_tables(c, &d->tors, 4, &c->tt, &c->ntt, &c->itort, &c->tors_lab, 1, 0, 1) ||
        setup_tables(c, &d->nb, 2, &c->nt, &c->nnbt, &c->inbondt, &c->nb_lab, 1, 0, 1)) {
        snprintf(err, (size_t)errlen, "%s", c->err);
        ox_destroy(c);
        return NULL;
    }
    /* OOP: IOOPT is filled in the given order only (RDINTERACT.f90:430-432) */
    if (setup_tables(c, &d->oop, 4, &c->ot, &c->nopt, &c->ioopt, &c->oop_lab, 1, 1, 0)) {
        snprintf(err, (size_t)errlen, "%s", c->err);
        ox_destroy(c);
        return NULL;
    }
    ftable(c);

    /* virtual sites (RDVIRTUAL.f90:40-140) */
    if (c->hybrid) {
        c->vs_type = (int *)malloc(sizeof(int) * (size_t)c->nvs);
        c->vs_center = (int *)malloc(sizeof(int) * (size_t)c->nvs);
        c->vs_off = (int *)malloc(sizeof(int) * (size_t)(c->nvs + 1));
        memcpy(c->vs_type, d->vs_type, sizeof(int) * (size_t)c->nvs);
        memcpy(c->vs_center, d->vs_center, sizeof(int) * (size_t)c->nvs);
        memcpy(c->vs_off, d->vs_off, sizeof(int) * (size_t)(c->nvs + 1));
        c->vs_atoms = (int *)malloc(sizeof(int) * (size_t)(d->vs_off[c->nvs] + 1));
        memcpy(c->vs_atoms, d->vs_atoms, sizeof(int) * (size_t)d->vs_off[c->nvs]);
        c->vs_invmass = (double *)calloc((size_t)c->ntype, sizeof(double));
        c->vs_coeff = (double *)calloc((size_t)c->ntype * c->ntype, sizeof(double));
        for (int i = 0; i < c->nvs; ++i) {
            int pos = c->natoms + i, vt = c->vs_type[i];
            c->itype[pos] = vt;
            c->type_label[pos] = 3;
            double m = 0.0;
            for (int p = c->vs_off[i]; p < c->vs_off[i + 1]; ++p) m += c->mass[c->itype[c->vs_atoms[p]]];
            c->vs_invmass[vt] = 1.0 / m;
            for (int p = c->vs_off[i]; p < c->vs_off[i + 1]; ++p) {
                int at = c->itype[c->vs_atoms[p]];
                c->vs_coeff[vt * c->ntype + at] = c->mass[at] / m;
            }
        }
    }

    /* ATOM / BEAD index lists (MAKE_LISTS.f90:13-60; main.F90:135-141 when not hybrid) */
    c->atom_list = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    c->bead_list = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    for (int i = 0; i < c->natoms; ++i) {
        if (c->type_label[i] == 1) c->atom_list[c->numatoms++] = i;
        else c->bead_list[c->numbeads++] = i;
    }
    c->ncoarse = c->numbeads;
    for (int i = 0; i < c->nvs; ++i) c->bead_list[c->ncoarse++] = c->natoms + i;
    c->ngroup = c->hybrid ? 2 : 1;
    c->g[0].n = c->numatoms; c->g[0].idx = c->atom_list;
    c->g[0].rcut = d->rcut_atom; c->g[0].rlist = d->rlist_atom; c->g[0].exclusion = d->nonbexc_atom;
    c->g[1].n = c->ncoarse; c->g[1].idx = c->bead_list;
    c->g[1].rcut = d->rcut_bead; c->g[1].rlist = d->rlist_bead; c->g[1].exclusion = d->nonbexc_bead;
    if (!c->hybrid) c->g[1].n = 0;

    /* topology */
    if (setlis(c, d)) {
        snprintf(err, (size_t)errlen, "%s", c->err);
        ox_destroy(c);
        return NULL;
    }
    build_connectivity(c, d);

    /* initial kinetic energy (main.F90:173-195) */
    c->mktemp = 2.0 / (3.0 * (c->natoms - 1));
    c->mktemp_atom = 2.0 / (3.0 * (c->numatoms - 1));
    c->mktemp_bead = 2.0 / (3.0 * (c->numbeads - 1));
    double ek1 = 0.0, ek2 = 0.0;
    for (int a = 0; a < c->numatoms; ++a) {
        int i = c->atom_list[a];
        const double *vi = c->v + 3 * i;
        ek1 += c->mass[c->itype[i]] * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
    }
    for (int a = 0; a < c->numbeads; ++a) {
        int i = c->bead_list[a];
        const double *vi = c->v + 3 * i;
        ek2 += c->mass[c->itype[i]] * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
    }
    c->ek[0] = 0.5 * ek1; c->ek[1] = 0.5 * ek2;
    c->temp = (c->ek[0] + c->ek[1]) * c->mktemp;
    c->temp_atom = c->ek[0] * c->mktemp_atom;
    c->temp_bead = c->ek[1] * c->mktemp_bead;
    if (c->ncoarse > 0) c->mts_f = (double *)calloc((size_t)9 * c->ncoarse, sizeof(double));

    /* main.F90:148-234: VS placement, fold, first list */
    if (c->hybrid) ox_virtual_def(c);
    ox_shift(c);
    if (ox_update_lists(c, 0)) {
        snprintf(err, (size_t)errlen, "%s", c->err);
        ox_destroy(c);
        return NULL;
    }
    return c;
}

static void free_tabs(tab_t *t, int n) {
    if (!t) return;
    for (int i = 0; i < n; ++i) { free(t[i].x); free(t[i].V); free(t[i].F); }
    free(t);
}

static void free_terms(terms_t *T) { free(T->off); free(T->a); free(T->b); free(T->c); }

void ox_destroy(ox_ctx *c) {
    if (!c) return;
    free(c->mass); free(c->invmass); free(c->tlabel); free(c->itype); free(c->type_label);
    free(c->nbonds); free(c->jbond); free(c->s); free(c->r); free(c->v); free(c->vt); free(c->f);
    free(c->fnb);
    free_tabs(c->bt, c->nbt); free_tabs(c->at, c->nat); free_tabs(c->tt, c->ntt);
    free_tabs(c->nt, c->nnbt); free_tabs(c->ot, c->nopt);
    free(c->ibondt); free(c->iangt); free(c->itort); free(c->inbondt); free(c->ioopt);
    free(c->bond_lab); free(c->angle_lab); free(c->tors_lab); free(c->oop_lab); free(c->nb_lab);
    free_terms(&c->ang); free_terms(&c->noang); free_terms(&c->tor); free_terms(&c->ftor);
    free_terms(&c->oop);
    free(c->conn_off); free(c->conn);
    for (int g = 0; g < 2; ++g) { free(c->g[g].nneb); free(c->g[g].off); free(c->g[g].nb); }
    free(c->atom_list); free(c->bead_list);
    free(c->vs_type); free(c->vs_center); free(c->vs_off); free(c->vs_atoms);
    free(c->vs_invmass); free(c->vs_coeff); free(c->mts_f); free(c->respa_f);
    free(c);
}

/* SHIFT.f90:20-48: fold the unwrapped coordinate into [-L/2, L/2] when any
 * component lies outside, using the reference's INT-based formula. */
static void fold(const double *s, double *r, const double *box2) {
    if (s[0] > box2[0] || s[0] < -box2[0] || s[1] > box2[1] || s[1] < -box2[1] ||
        s[2] > box2[2] || s[2] < -box2[2]) {
        for (int k = 0; k < 3; ++k) {
            double x = s[k] / box2[k];
            double sp = x - 2.0 * trunc(x / 2.0);
            r[k] = (sp - 2.0 * trunc(sp)) * box2[k];
        }
    } else {
        r[0] = s[0]; r[1] = s[1]; r[2] = s[2];
    }
}

void ox_shift(ox_ctx *c) {
    for (int i = 0; i < c->natoms; ++i) fold(c->s + 3 * i, c->r + 3 * i, c->box2);
}

/* VIRTUAL_DEF.f90:8-50 (COM accumulators zeroed in all three components). */
void ox_virtual_def(ox_ctx *c) {
    for (int i = 0; i < c->nvs; ++i) {
        int pos = c->natoms + i;
        double *sp = c->s + 3 * pos;
        if (c->vs_center[i] >= 0) {
            const double *sj = c->s + 3 * c->vs_center[i];
            sp[0] = sj[0]; sp[1] = sj[1]; sp[2] = sj[2];
        } else {
            double sum[3] = {0, 0, 0};
            for (int p = c->vs_off[i]; p < c->vs_off[i + 1]; ++p) {
                int k = c->vs_atoms[p];
                double m = c->mass[c->itype[k]];
                for (int d = 0; d < 3; ++d) sum[d] += m * c->s[3 * k + d];
            }
            double im = c->vs_invmass[c->vs_type[i]];
            for (int d = 0; d < 3; ++d) sp[d] = sum[d] * im;
        }
        fold(sp, c->r + 3 * pos, c->box2);
    }
}

static int excluded(const ox_ctx *c, int I, int J) {
    for (int p = c->conn_off[I]; p < c->conn_off[I + 1]; ++p)
        if (c->conn[p] == J) return 1;
    if (c->hybrid && I >= c->natoms && J >= c->natoms) return 1;   /* VS-VS */
    return 0;
}

/* pair criterion of NEW_NEIGHBOUR_WITHLIST.f90:64-69 */
static double pair_r2(const ox_ctx *c, int I, int J) {
    double r2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        double d = c->r[3 * I + k] - c->r[3 * J + k];
        d = d - anint(d * c->boxinv[k]) * c->box[k];
        r2 += d * d;
    }
    return r2;
}

/* UPDATE_NEIGHBOURLIST.F90:53-85 -> LINKS.f90:5-25, MAPS.f90:20-77,
 * NEW_NEIGHBOUR_WITHLIST.f90:5-127 for each group. */
static int build_group(ox_ctx *c, group_t *g, int mode) {
    free(g->nneb); free(g->off); free(g->nb);
    g->nneb = (int *)calloc((size_t)g->n + 1, sizeof(int));
    g->off = (int *)calloc((size_t)g->n + 1, sizeof(int));
    ivec out = {0};
    double rl2 = g->rlist * g->rlist;
    int nc[3];
    for (int k = 0; k < 3; ++k) nc[k] = (int)(c->box[k] / g->rlist);
    int small = nc[0] < 3 || nc[1] < 3 || nc[2] < 3;
    if (mode == 1 || small) {
        for (int A = 0; A < g->n; ++A) {
            int I = g->idx[A];
            g->off[A] = out.n;
            for (int B = A + 1; B < g->n; ++B) {
                int J = g->idx[B];
                if (excluded(c, I, J)) continue;
                if (pair_r2(c, I, J) < rl2) iv_push(&out, J);
            }
            g->nneb[A] = out.n - g->off[A];
        }
    } else {
        int ncell = nc[0] * nc[1] * nc[2];
        int *head = (int *)malloc(sizeof(int) * (size_t)ncell);
        int *lcl = (int *)malloc(sizeof(int) * (size_t)g->n);
        int *cell = (int *)malloc(sizeof(int) * (size_t)g->n);
        int *map = (int *)malloc(sizeof(int) * (size_t)ncell * 13);
        for (int q = 0; q < ncell; ++q) head[q] = -1;
        for (int A = 0; A < g->n; ++A) {
            int I = g->idx[A], ci[3];
            for (int k = 0; k < 3; ++k) {
                ci[k] = (int)((c->r[3 * I + k] * c->boxinv[k] + 0.5) * nc[k]);
                if (ci[k] >= nc[k]) ci[k] = nc[k] - 1;
                if (ci[k] < 0) ci[k] = 0;
            }
            int jc = ci[0] * nc[2] * nc[1] + ci[1] * nc[2] + ci[2];
            lcl[A] = head[jc];
            head[jc] = A;
            cell[A] = jc;
        }
        static const int st[13][3] = {{1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {-1, 1, 0}, {1, 0, -1}, {1, 1, -1},
                                      {0, 1, -1}, {-1, 1, -1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}, {-1, 1, 1},
                                      {0, 0, 1}};
        for (int ix = 0; ix < nc[0]; ++ix)
            for (int iy = 0; iy < nc[1]; ++iy)
                for (int iz = 0; iz < nc[2]; ++iz) {
                    int base = (ix * nc[1] * nc[2] + iy * nc[2] + iz) * 13;
                    for (int q = 0; q < 13; ++q) {
                        int x = (ix + st[q][0] + nc[0]) % nc[0];
                        int y = (iy + st[q][1] + nc[1]) % nc[1];
                        int z = (iz + st[q][2] + nc[2]) % nc[2];
                        map[base + q] = z + y * nc[2] + x * nc[2] * nc[1];
                    }
                }
        for (int A = 0; A < g->n; ++A) {
            int I = g->idx[A];
            g->off[A] = out.n;
            int cnt = 0;
            for (int B = lcl[A]; B >= 0; B = lcl[B]) {
                int J = g->idx[B];
                if (excluded(c, I, J)) continue;
                if (pair_r2(c, I, J) < rl2) {
                    iv_push(&out, J);
                    if (++cnt == MAXNAB) { snprintf(c->err, sizeof c->err, "LIST TOO SMALL"); return -1; }
                }
            }
            for (int q = 0; q < 13; ++q) {
                for (int B = head[map[cell[A] * 13 + q]]; B >= 0; B = lcl[B]) {
                    int J = g->idx[B];
                    if (excluded(c, I, J)) continue;
                    if (pair_r2(c, I, J) < rl2) {
                        iv_push(&out, J);
                        if (++cnt == MAXNAB) { snprintf(c->err, sizeof c->err, "LIST TOO SMALL"); return -1; }
                    }
                }
            }
            g->nneb[A] = cnt;
        }
        free(head); free(lcl); free(cell); free(map);
    }
    g->off[g->n] = out.n;
    g->nb = out.data;
    g->nnb = out.n;
    return 0;
}

int ox_update_lists(ox_ctx *c, int mode) {
    for (int g = 0; g < c->ngroup; ++g)
        if (build_group(c, &c->g[g], mode)) return -1;
    return 0;
}

static int cmp_pair(const void *a, const void *b) {
    const int *x = (const int *)a, *y = (const int *)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

int64_t ox_pairs(ox_ctx *c, int group, int *out, int64_t cap) {
    if (group >= c->ngroup) return 0;
    group_t *g = &c->g[group];
    int64_t n = g->nnb;
    if (!out || cap < n) return n;
    int64_t p = 0;
    for (int A = 0; A < g->n; ++A) {
        int I = g->idx[A];
        for (int q = g->off[A]; q < g->off[A] + g->nneb[A]; ++q) {
            int J = g->nb[q];
            out[2 * p] = I < J ? I : J;
            out[2 * p + 1] = I < J ? J : I;
            ++p;
        }
    }
    qsort(out, (size_t)n, 2 * sizeof(int), cmp_pair);
    return n;
}

int ox_exclusions(ox_ctx *c, int *off, int *idx, int cap) {
    int n = c->conn_off[c->nitems];
    if (!off || cap < n) return n;
    memcpy(off, c->conn_off, sizeof(int) * (size_t)(c->nitems + 1));
    memcpy(idx, c->conn, sizeof(int) * (size_t)n);
    return n;
}

int ox_terms(ox_ctx *c, int kind, int *out, int cap) {
    terms_t *T = kind == 0 ? &c->ang : kind == 1 ? &c->tor : kind == 2 ? &c->noang
               : kind == 3 ? &c->ftor : &c->oop;
    int w = (kind == 0 || kind == 2) ? 3 : 4;
    int n = T->off[c->nitems];
    if (!out || cap < n) return n;
    int p = 0;
    for (int I = 0; I < c->nitems; ++I)
        for (int q = T->off[I]; q < T->off[I + 1]; ++q) {
            out[w * p] = I;
            out[w * p + 1] = T->a[q];
            out[w * p + 2] = T->b[q];
            if (w == 4) out[w * p + 3] = T->c[q];
            ++p;
        }
    return n;
}

/* NONBONDED_FORCE.F90:11-109 */
static void nonbonded(ox_ctx *c, group_t *g) {
    double rc2 = g->rcut * g->rcut;
    for (int A = 0; A < g->n; ++A) {
        int I = g->idx[A], TI = c->itype[I];
        double fi[3] = {0, 0, 0};
        for (int q = g->off[A]; q < g->off[A] + g->nneb[A]; ++q) {
            int J = g->nb[q], TJ = c->itype[J];
            int tij = c->inbondt[ti2(c, TI, TJ)];
            double d[3];
            for (int k = 0; k < 3; ++k) {
                d[k] = c->r[3 * I + k] - c->r[3 * J + k];
                d[k] = d[k] - anint(d[k] * c->boxinv[k]) * c->box[k];
            }
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(r2 < rc2)) continue;
            const tab_t *t = &c->nt[tij];
            double rij = sqrt(r2);
            int ni = (int)(rij / t->bin);
            if (ni + 1 > t->ndat + 1) ni = t->ndat;   /* beyond the table: zero row */
            double alpha = (rij - t->x[ni]) / t->bin;
            double fij = t->F[ni] * (1.0 - alpha) + t->F[ni + 1] * alpha;
            double vij = t->V[ni] * (1.0 - alpha) + t->V[ni + 1] * alpha;
            int li = c->type_label[I], lj = c->type_label[J];
            if (li == 1) c->vnb[0] += vij;
            else if (li == 2) { if (lj == 2) c->vnb[1] += vij; else if (lj == 3) c->vnb[2] += vij; }
            else if (li == 3) c->vnb[2] += vij;
            for (int k = 0; k < 3; ++k) {
                fi[k] += fij * d[k];
                c->f[3 * J + k] -= fij * d[k];
            }
            c->pt[0] += fij * d[0] * d[0];
            c->pt[1] += fij * d[1] * d[1];
            c->pt[2] += fij * d[2] * d[2];
            c->pt[3] += fij * d[1] * d[0];
            c->pt[4] += fij * d[2] * d[0];
            c->pt[5] += fij * d[2] * d[1];
        }
        for (int k = 0; k < 3; ++k) c->f[3 * I + k] += fi[k];
    }
}

/* DISTRIBUTE_VSFORCE.f90:12-41 */
static void distribute_vsforce(ox_ctx *c) {
    for (int i = 0; i < c->nvs; ++i) {
        int pos = c->natoms + i, vt = c->vs_type[i];
        for (int p = c->vs_off[i]; p < c->vs_off[i + 1]; ++p) {
            int j = c->vs_atoms[p];
            double cf = c->vs_coeff[vt * c->ntype + c->itype[j]];
            for (int k = 0; k < 3; ++k) c->f[3 * j + k] += c->f[3 * pos + k] * cf;
        }
        c->f[3 * pos] = c->f[3 * pos + 1] = c->f[3 * pos + 2] = 0.0;
    }
}

static double sq3(const double *a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

/* BONDED_FORCE.F90:13-105 with BONDS.inc, ANGLES.inc, TORSIONS.inc, OUTOFPLANES.inc */
static int bonded(ox_ctx *c) {
    const double R2D = 180.0 / PI;
    double *fl = (double *)calloc((size_t)c->nitems * 3, sizeof(double));
    for (int I = 0; I < c->nitems; ++I) {
        const double *ri = c->s + 3 * I;
        int TI = c->itype[I];
        double fi[3] = {0, 0, 0};
        /* BONDS.inc:3-57 */
        if (I < c->natoms)
            for (int a = 0; a < c->nbonds[I]; ++a) {
                int J = c->jbond[4 * I + a];
                if (J <= I) continue;
                int tij = c->ibondt[ti2(c, TI, c->itype[J])];
                const tab_t *t = &c->bt[tij];
                double d[3];
                for (int k = 0; k < 3; ++k) d[k] = ri[k] - c->s[3 * J + k];
                double rij = sqrt(sq3(d));
                int ni = (int)((rij - t->x[0]) / t->bin);
                if (ni > t->ndat) { ni = t->ndat; c->overext++; }
                else if (ni < 0) ni = 0;
                double alpha = (rij - t->x[ni]) / t->bin;
                double fij = t->F[ni] * (1.0 - alpha) + alpha * t->F[ni + 1];
                double vij = t->V[ni] * (1.0 - alpha) + alpha * t->V[ni + 1];
                c->vbond[c->bond_lab[tij] - 1] += vij;
                for (int k = 0; k < 3; ++k) {
                    fi[k] += fij * d[k];
                    fl[3 * J + k] -= fij * d[k];
                }
            }
        /* ANGLES.inc:2-75 */
        for (int p = c->ang.off[I]; p < c->ang.off[I + 1]; ++p) {
            int J = c->ang.a[p], K = c->ang.b[p];
            int tijk = c->iangt[ti3(c, TI, c->itype[J], c->itype[K])];
            const tab_t *t = &c->at[tijk];
            double ij[3], kj[3], M[3], N[3], S[3];
            for (int k = 0; k < 3; ++k) { ij[k] = ri[k] - c->s[3 * J + k]; kj[k] = c->s[3 * K + k] - c->s[3 * J + k]; }
            double rij = sqrt(sq3(ij)), rkj = sqrt(sq3(kj));
            double ct = (ij[0] * kj[0] + ij[1] * kj[1] + ij[2] * kj[2]) / rij / rkj;
            if (ct > 1.0) ct = 1.0;   /* acos domain: FP64 rounding only */
            if (ct < -1.0) ct = -1.0;
            double theta = acos(ct) * R2D;
            M[0] = -(ij[1] * kj[2] - ij[2] * kj[1]);
            M[1] = -(-ij[0] * kj[2] + ij[2] * kj[0]);
            M[2] = -(ij[0] * kj[1] - ij[1] * kj[0]);
            double rm = sqrt(sq3(M));
            for (int k = 0; k < 3; ++k) M[k] /= rm;
            N[0] = (M[1] * ij[2] - M[2] * ij[1]) / rij;
            N[1] = (-M[0] * ij[2] + M[2] * ij[0]) / rij;
            N[2] = (M[0] * ij[1] - M[1] * ij[0]) / rij;
            S[0] = -(M[1] * kj[2] - M[2] * kj[1]) / rkj;
            S[1] = -(-M[0] * kj[2] + M[2] * kj[0]) / rkj;
            S[2] = -(M[0] * kj[1] - M[1] * kj[0]) / rkj;
            int ni = (int)(theta / t->bin);
            double fijk, vijk;
            if (theta == 180.0) { fijk = t->F[ni]; vijk = t->V[ni]; }
            else {
                double alpha = (theta - t->x[ni]) / t->bin;
                fijk = t->F[ni] * (1.0 - alpha) + alpha * t->F[ni + 1];
                vijk = t->V[ni] * (1.0 - alpha) + alpha * t->V[ni + 1];
            }
            c->vang[c->angle_lab[tijk] - 1] += vijk;
            for (int k = 0; k < 3; ++k) {
                double fa = fijk * N[k] / rij, fk = fijk * S[k] / rkj;
                fi[k] += fa;
                fl[3 * K + k] += fk;
                fl[3 * J + k] -= fk + fa;
            }
        }
        /* TORSIONS.inc:1-152 */
        for (int p = c->tor.off[I]; p < c->tor.off[I + 1]; ++p) {
            int J = c->tor.a[p], K = c->tor.b[p], L = c->tor.c[p];
            int tt = c->itort[ti4(c, TI, c->itype[J], c->itype[K], c->itype[L])];
            const tab_t *t = &c->tt[tt];
            double ij[3], jk[3], kl[3], M[3], N[3];
            for (int k = 0; k < 3; ++k) {
                ij[k] = ri[k] - c->s[3 * J + k];
                jk[k] = c->s[3 * J + k] - c->s[3 * K + k];
                kl[k] = c->s[3 * K + k] - c->s[3 * L + k];
            }
            double rij = sqrt(sq3(ij)), rjk = sqrt(sq3(jk)), rkl = sqrt(sq3(kl));
            double cosm, sinm, cotm, cosn, sinn, cotn;
            M[0] = ij[1] * jk[2] - ij[2] * jk[1];
            M[1] = -ij[0] * jk[2] + ij[2] * jk[0];
            M[2] = ij[0] * jk[1] - ij[1] * jk[0];
            double rm = sqrt(sq3(M));
            if (rm == 0.0) {   /* TORSIONS.inc:47-59: perturb x of atom i */
                ij[0] = ri[0] * 0.99 - c->s[3 * J];
                rij = sqrt(sq3(ij));
                M[0] = ij[1] * jk[2] - ij[2] * jk[1];
                M[1] = -ij[0] * jk[2] + ij[2] * jk[0];
                M[2] = ij[0] * jk[1] - ij[1] * jk[0];
                rm = sqrt(sq3(M));
            }
            for (int k = 0; k < 3; ++k) M[k] /= rm;
            cosm = (ij[0] * jk[0] + ij[1] * jk[1] + ij[2] * jk[2]) / rij / rjk;
            sinm = sqrt(1 - cosm * cosm);
            cotm = cosm / sinm;
            N[0] = jk[1] * kl[2] - jk[2] * kl[1];
            N[1] = -jk[0] * kl[2] + jk[2] * kl[0];
            N[2] = jk[0] * kl[1] - jk[1] * kl[0];
            double rn = sqrt(sq3(N));
            if (rn == 0.0) {   /* TORSIONS.inc:75-87: perturb x of atom l */
                kl[0] = c->s[3 * K] - c->s[3 * L] * 0.99;
                rkl = sqrt(sq3(kl));
                N[0] = jk[1] * kl[2] - jk[2] * kl[1];
                N[1] = -jk[0] * kl[2] + jk[2] * kl[0];
                N[2] = jk[0] * kl[1] - jk[1] * kl[0];
                rn = sqrt(sq3(N));
            }
            for (int k = 0; k < 3; ++k) N[k] /= rn;
            cosn = (jk[0] * kl[0] + jk[1] * kl[1] + jk[2] * kl[2]) / rjk / rkl;
            sinn = sqrt(1 - cosn * cosn);
            cotn = cosn / sinn;
            double cost = M[0] * N[0] + M[1] * N[1] + M[2] * N[2];
            double signt = -(M[0] * kl[0] + M[1] * kl[1] + M[2] * kl[2]);
            if (cost > 1.0) cost = 1.0;
            else if (cost < -1.0) cost = -1.0;
            double phi = acos(cost) * R2D;
            if (signt < 0) phi = 360.0 - phi;
            if (!(phi >= 0.0 && phi <= 360.0)) {   /* NaN: collinear atoms the 0.99 perturbation left collinear */
                snprintf(c->err, sizeof c->err, "torsion %d-%d-%d-%d undefined", I + 1, J + 1, K + 1, L + 1);
                free(fl);
                return -1;
            }
            int ni = (int)(phi / t->bin);
            if (ni > t->ndat) { snprintf(c->err, sizeof c->err, "torsion table %d overflow", tt + 1); free(fl); return -1; }
            double fijkl, vijkl;
            if (phi == 360.0) { fijkl = t->F[ni]; vijkl = t->V[ni]; }
            else {
                double alpha = (phi - t->x[ni]) / t->bin;
                fijkl = t->F[ni] * (1.0 - alpha) + alpha * t->F[ni + 1];
                vijkl = t->V[ni] * (1.0 - alpha) + alpha * t->V[ni + 1];
            }
            c->vtor[c->tors_lab[tt] - 1] += vijkl;
            for (int k = 0; k < 3; ++k) {
                double f1 = -fijkl * M[k] / sinm / rij;
                double f4 = fijkl * N[k] / sinn / rkl;
                double f12 = fijkl * (cotm * M[k] + cotn * N[k]) / rjk;
                fi[k] += f1;
                fl[3 * L + k] += f4;
                fl[3 * J + k] += -f1 + f12;
                fl[3 * K + k] += -f4 - f12;
            }
        }
        /* OUTOFPLANES.inc:1-141 */
        for (int p = c->oop.off[I]; p < c->oop.off[I + 1]; ++p) {
            int J = c->oop.a[p], K = c->oop.b[p], L = c->oop.c[p];
            int to = c->ioopt[ti4(c, TI, c->itype[J], c->itype[K], c->itype[L])];
            const tab_t *t = &c->ot[to];
            double ij[3], jk[3], kl[3], M[3], N[3], mN[3];
            for (int k = 0; k < 3; ++k) {
                ij[k] = ri[k] - c->s[3 * J + k];
                jk[k] = c->s[3 * K + k] - c->s[3 * J + k];
                kl[k] = c->s[3 * K + k] - c->s[3 * L + k];
            }
            double rij = sqrt(sq3(ij)), rjk = sqrt(sq3(jk)), rkl = sqrt(sq3(kl));
            M[0] = ij[1] * jk[2] - ij[2] * jk[1];
            M[1] = -ij[0] * jk[2] + ij[2] * jk[0];
            M[2] = ij[0] * jk[1] - ij[1] * jk[0];
            double rm = sqrt(sq3(M));
            for (int k = 0; k < 3; ++k) M[k] /= rm;
            double cosm = (ij[0] * jk[0] + ij[1] * jk[1] + ij[2] * jk[2]) / rij / rjk;
            double sinm = sqrt(1 - cosm * cosm), cotm = cosm / sinm;
            N[0] = jk[1] * kl[2] - jk[2] * kl[1];
            N[1] = -jk[0] * kl[2] + jk[2] * kl[0];
            N[2] = jk[0] * kl[1] - jk[1] * kl[0];
            double rn = sqrt(sq3(N));
            for (int k = 0; k < 3; ++k) N[k] /= rn;
            double cosn = (jk[0] * kl[0] + jk[1] * kl[1] + jk[2] * kl[2]) / rjk / rkl;
            double sinn = sqrt(1 - cosn * cosn), cotn = cosn / sinn;
            double cost = M[0] * N[0] + M[1] * N[1] + M[2] * N[2];
            mN[0] = M[1] * N[2] - M[2] * N[1];
            mN[1] = -M[0] * N[2] + M[2] * N[0];
            mN[2] = M[0] * N[1] - M[1] * N[0];
            double signt = copysign(1.0, mN[0] * jk[0] + mN[1] * jk[1] + mN[2] * jk[2]);
            if (cost > 1.0 + 1e-7 || cost < -(1.0 + 1e-7)) {
                snprintf(c->err, sizeof c->err, "OOP cos out of range"); free(fl); return -1;
            }
            if (cost > 1.0) cost = 1.0;
            else if (cost < -1.0) cost = -1.0;
            double phi = signt * acos(cost) * R2D;
            double phit = 180.0 + phi;
            int ni = (int)anint(phit / t->bin);
            if (ni < 0 || ni > t->ndat) { snprintf(c->err, sizeof c->err, "OOP table overflow"); free(fl); return -1; }
            double alpha = fabs(fabs(phi) - fabs(t->x[ni])) / t->bin;
            if (alpha > 1.0) { ni = ni + 1; alpha = fabs(fabs(phi) - fabs(t->x[ni])) / t->bin; }
            double fijkl = t->F[ni] * (1.0 - alpha) + alpha * t->F[ni + 1];
            double vijkl = t->V[ni] * (1.0 - alpha) + alpha * t->V[ni + 1];
            c->voop[c->oop_lab[to] - 1] += vijkl;
            for (int k = 0; k < 3; ++k) {
                double f1 = fijkl * M[k] / sinm / rij;
                double f4 = -fijkl * N[k] / sinn / rkl;
                double f12 = fijkl * (cotm * M[k] + cotn * N[k]) / rjk;
                fi[k] += f1;
                fl[3 * L + k] += f4;
                fl[3 * J + k] += -f1 + f12;
                fl[3 * K + k] += -f4 - f12;
            }
        }
        for (int k = 0; k < 3; ++k) fl[3 * I + k] += fi[k];
    }
    for (int i = 0; i < 3 * c->nitems; ++i) c->f[i] += fl[i];
    free(fl);
    return 0;
}

static void zero_force_obs(ox_ctx *c) {
    memset(c->f, 0, sizeof(double) * 3 * (size_t)c->nitems);
    memset(c->fnb, 0, sizeof(double) * 3 * (size_t)c->natoms);
    for (int k = 0; k < 6; ++k) c->pt[k] = 0.0;
    for (int k = 0; k < 3; ++k) c->vnb[k] = c->vbond[k] = c->vang[k] = c->vtor[k] = c->voop[k] = 0.0;
}

static int finish_force(ox_ctx *c) {
    for (int i = 0; i < 3 * c->natoms; ++i) c->fnb[i] = c->f[i];
    if (bonded(c)) return -1;
    if (c->hybrid) distribute_vsforce(c);
    /* NEW_FORCE.F90:73-80: bonded virial over unwrapped coordinates */
    for (int i = 0; i < c->natoms; ++i) {
        const double *f = c->f + 3 * i, *fn = c->fnb + 3 * i, *s = c->s + 3 * i;
        c->pt[0] += (f[0] - fn[0]) * s[0];
        c->pt[1] += (f[1] - fn[1]) * s[1];
        c->pt[2] += (f[2] - fn[2]) * s[2];
        c->pt[3] += (f[1] - fn[1]) * s[0];
        c->pt[4] += (f[2] - fn[2]) * s[0];
        c->pt[5] += (f[2] - fn[2]) * s[1];
    }
    return 0;
}

/* NEW_FORCE.F90:7-83 */
void ox_new_force(ox_ctx *c) {
    zero_force_obs(c);
    nonbonded(c, &c->g[0]);
    if (c->hybrid) {
        nonbonded(c, &c->g[1]);
        distribute_vsforce(c);
    }
    finish_force(c);
}

/* MTS_LOOP.f90:71-224 (MTS_FORCE with MTS_SAVEFORCE/APPROX/SAVEAUXS/LOADAUXS) */
static int mts_force(ox_ctx *c, int step) {
    zero_force_obs(c);
    int mod = step % (3 + c->nmts);
    if (mod == 1 || mod == 2 || mod == 3) {
        nonbonded(c, &c->g[1]);
        for (int a = 0; a < c->ncoarse; ++a) {
            int id = c->bead_list[a];
            for (int k = 0; k < 3; ++k) c->mts_f[((mod - 1) * c->ncoarse + a) * 3 + k] = c->f[3 * id + k];
        }
        for (int k = 0; k < 3; ++k) c->mts_vnb[mod - 1][k] = c->vnb[k];
        /* MTS_SAVEAUXS never stores PT23 (MTS_LOOP.f90:193-199) */
        for (int k = 0; k < 5; ++k) c->mts_pt[mod - 1][k] = c->pt[k];
    } else {
        double I = (double)mod;
        if (mod == 0 && !c->mts_quirk) I = 3.0 + c->nmts;
        for (int a = 0; a < c->ncoarse; ++a) {
            int id = c->bead_list[a];
            for (int k = 0; k < 3; ++k) {
                double f3 = c->mts_f[(2 * c->ncoarse + a) * 3 + k], f1 = c->mts_f[a * 3 + k];
                c->f[3 * id + k] = f3 + (I - 3.0) * 0.5 * (f3 - f1);
            }
        }
        for (int k = 0; k < 3; ++k)
            c->vnb[k] = c->mts_vnb[2][k] + (I - 3.0) * 0.5 * (c->mts_vnb[2][k] - c->mts_vnb[0][k]);
        for (int k = 0; k < 6; ++k)
            c->pt[k] = c->mts_pt[2][k] + (I - 3.0) * 0.5 * (c->mts_pt[2][k] - c->mts_pt[0][k]);
    }
    distribute_vsforce(c);
    nonbonded(c, &c->g[0]);
    return finish_force(c);
}

/* MOVE.F90:9-118: leap-frog with a Berendsen factor per sub-population. */
void ox_move(ox_ctx *c) {
    double lam[2] = {1.0, 1.0};
    if (c->ensemble == 1 || c->ensemble == 2) {
        lam[0] = sqrt(1.0 + c->fac * ((c->temp_in / c->temp_atom) - 1.0));
        lam[1] = sqrt(1.0 + c->fac * ((c->temp_in / c->temp_bead) - 1.0));
    }
    double pk[2][6] = {{0}};
    for (int pop = 0; pop < 2; ++pop) {
        int n = pop == 0 ? c->numatoms : c->numbeads;
        const int *lst = pop == 0 ? c->atom_list : c->bead_list;
        for (int a = 0; a < n; ++a) {
            int i = lst[a];
            double m = c->mass[c->itype[i]], im = c->invmass[c->itype[i]];
            double vt[3];
            for (int k = 0; k < 3; ++k) {
                double v0 = c->v[3 * i + k];
                double v1 = (v0 + c->dt * c->f[3 * i + k] * im) * lam[pop];
                c->v[3 * i + k] = v1;
                c->s[3 * i + k] += c->dt * v1;
                vt[k] = 0.5 * (v1 + v0);
                c->vt[3 * i + k] = vt[k];
            }
            pk[pop][0] += m * vt[0] * vt[0];
            pk[pop][1] += m * vt[1] * vt[1];
            pk[pop][2] += m * vt[2] * vt[2];
            pk[pop][3] += m * vt[0] * vt[1];
            pk[pop][4] += m * vt[0] * vt[2];
            pk[pop][5] += m * vt[1] * vt[2];
        }
    }
    for (int k = 0; k < 6; ++k) c->pt[k] += pk[0][k] + pk[1][k];
    c->ek[0] = 0.5 * (pk[0][0] + pk[0][1] + pk[0][2]);
    c->ek[1] = 0.5 * (pk[1][0] + pk[1][1] + pk[1][2]);
    c->temp = (c->ek[0] + c->ek[1]) * c->mktemp;
    c->temp_atom = c->ek[0] * c->mktemp_atom;
    c->temp_bead = c->ek[1] * c->mktemp_bead;
    double vol = c->box[0] * c->box[1] * c->box[2];
    for (int k = 0; k < 6; ++k) c->pt[k] *= 1.0 / vol;
}

/* MOMENTUM.f90:5-32 */
void ox_momentum(ox_ctx *c) {
    double sum[3] = {0, 0, 0};
    for (int i = 0; i < c->natoms; ++i)
        for (int k = 0; k < 3; ++k) sum[k] += c->v[3 * i + k] * c->mass[c->itype[i]];
    for (int k = 0; k < 3; ++k) sum[k] /= (double)c->natoms;
    for (int i = 0; i < c->natoms; ++i)
        for (int k = 0; k < 3; ++k) c->v[3 * i + k] -= sum[k] * c->invmass[c->itype[i]];
}

/* Multiple time step for non-hybrid systems (r-RESPA in leap-frog form; no counterpart in the reference,
 * so no reference file:line): the nonbonded force is evaluated on steps 1, k+1, 2k+1, ... and applied with
 * weight k (the two half kicks of the slow force around an outer step boundary merged into one), the bonded
 * terms on every step.  The other steps report the last evaluation's nonbonded energy and virial.  With
 * k = 1 this is NEW_FORCE.  Pinned by k = 1 against the plain loop and by time reversibility. */
static int respa_force(ox_ctx *c, int step) {
    const int nb_step = (step - 1) % c->respa == 0;
    const size_t m = 3 * (size_t)c->nitems;
    zero_force_obs(c);
    if (nb_step) {
        nonbonded(c, &c->g[0]);
        memcpy(c->respa_f, c->f, sizeof(double) * m);
        for (int k = 0; k < 3; ++k) c->respa_vnb[k] = c->vnb[k];
        for (int k = 0; k < 6; ++k) c->respa_pt[k] = c->pt[k];
    } else {
        for (int k = 0; k < 3; ++k) c->vnb[k] = c->respa_vnb[k];
        for (int k = 0; k < 6; ++k) c->pt[k] = c->respa_pt[k];
    }
    if (finish_force(c)) return -1;
    if (nb_step)
        for (size_t q = 0; q < m; ++q) c->f[q] += (c->respa - 1) * c->respa_f[q];
    return 0;
}

int ox_set_respa(ox_ctx *c, int k) {
    if (k < 1 || (k > 1 && (c->hybrid || c->nmts > 0))) return -1;
    c->respa = k;
    if (!c->respa_f) c->respa_f = (double *)calloc(3 * (size_t)c->nitems, sizeof(double));
    return 0;
}

/* SCALEBP.f90:6-44: the instantaneous pressure (PT11 + PT22 + PT33)/3 after MOVE goes into a ring of LIMAVP
 * = 10 entries; every LIMAVP-th step its mean P scales the box and the positions of the atoms and beads by
 * MIU = (1 + (P - PRESSURE) PFAC)^(1/3) (virtual sites follow at the next VIRTUAL_DEF).  The cell grid follows
 * the box at the next list update (UPDATE_NEIGHBOURLIST.F90:14-48; build_group takes it from the box). */
void ox_scalebp(ox_ctx *c, int step) {
    c->iavp = c->iavp % 10 + 1;
    c->sp[c->iavp - 1] = (c->pt[0] + c->pt[1] + c->pt[2]) / 3.0;
    if (step % 10 != 0) return;
    double avp = 0.0;
    for (int k = 0; k < 10; ++k) avp += c->sp[k];
    const double p = avp / 10.0;
    const double miu = cbrt(1.0 + (p - c->pressure) * c->pfac);
    for (int k = 0; k < 3; ++k) {
        c->box[k] *= miu;
        c->boxinv[k] = 1.0 / c->box[k];
        c->box2[k] = c->box[k] / 2.0;
    }
    for (int i = 0; i < c->natoms; ++i)
        for (int k = 0; k < 3; ++k) c->s[3 * i + k] *= miu;
}

/* NEW_LOOP.F90:19-156 and MTS_LOOP.f90:13-60 */
int ox_run(ox_ctx *c, int nsteps, int step0) {
    for (int step = step0 + 1; step <= step0 + nsteps; ++step) {
        ox_shift(c);
        if (c->hybrid) ox_virtual_def(c);
        if (step % c->nupdate == 0)
            if (ox_update_lists(c, 0)) return -1;
        if (c->nmts > 0 && c->hybrid) {
            if (mts_force(c, step)) return -1;
        } else if (c->respa > 1) {
            if (respa_force(c, step)) return -1;
        } else {
            zero_force_obs(c);
            nonbonded(c, &c->g[0]);
            if (c->hybrid) { nonbonded(c, &c->g[1]); distribute_vsforce(c); }
            if (finish_force(c)) return -1;
        }
        ox_move(c);
        if (step % c->halt_drift == 0) ox_momentum(c);
        if (c->ensemble == 2) ox_scalebp(c, step);   /* NEW_LOOP.F90:120-122, MTS_LOOP.f90:42-44 */
    }
    return 0;
}

void ox_get_obs(ox_ctx *c, double *o) {
    for (int k = 0; k < 3; ++k) {
        o[k] = c->vnb[k]; o[3 + k] = c->vbond[k]; o[6 + k] = c->vang[k];
        o[9 + k] = c->vtor[k]; o[12 + k] = c->voop[k];
    }
    for (int k = 0; k < 6; ++k) o[15 + k] = c->pt[k];
    o[21] = c->ek[0]; o[22] = c->ek[1];
    o[23] = c->temp; o[24] = c->temp_atom; o[25] = c->temp_bead;
    o[26] = c->box[0] * c->box[1] * c->box[2];
    o[27] = (double)c->overext;
    for (int k = 0; k < 3; ++k) o[28 + k] = c->box[k];
}

int ox_nitems(ox_ctx *c) { return c->nitems; }

void ox_get_forces(ox_ctx *c, double *f) { memcpy(f, c->f, sizeof(double) * 3 * (size_t)c->nitems); }

void ox_get_state(ox_ctx *c, double *s, double *v, double *r) {
    if (s) memcpy(s, c->s, sizeof(double) * 3 * (size_t)c->nitems);
    if (v) memcpy(v, c->v, sizeof(double) * 3 * (size_t)c->natoms);
    if (r) memcpy(r, c->r, sizeof(double) * 3 * (size_t)c->nitems);
}

void ox_set_state(ox_ctx *c, const double *s, const double *v) {
    if (s) memcpy(c->s, s, sizeof(double) * 3 * (size_t)c->natoms);
    if (v) memcpy(c->v, v, sizeof(double) * 3 * (size_t)c->natoms);
    if (c->hybrid) ox_virtual_def(c);
    ox_shift(c);
}

int ox_table(ox_ctx *c, int kind, int t, double *x, double *V, double *F, int cap) {
    tab_t *tb = kind == 0 ? c->bt : kind == 1 ? c->at : kind == 2 ? c->tt : kind == 3 ? c->nt : c->ot;
    int nt = kind == 0 ? c->nbt : kind == 1 ? c->nat : kind == 2 ? c->ntt : kind == 3 ? c->nnbt : c->nopt;
    if (t < 0 || t >= nt) return -1;
    int n = tb[t].ndat + 2;
    if (!x || cap < n) return n;
    memcpy(x, tb[t].x, sizeof(double) * (size_t)n);
    memcpy(V, tb[t].V, sizeof(double) * (size_t)n);
    memcpy(F, tb[t].F, sizeof(double) * (size_t)n);
    return n;
}
