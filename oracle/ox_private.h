/*
 * Internal layout of the FP64 oracle's context (ibx_oracle.c), shared with the
 * FP32 OpenMP port (cpu_port.c), which takes its set-up (FTABLE'd tables,
 * SETLIS terms, CONNECTED_TO lists) from an oracle context.
 *
 * TEST INFRASTRUCTURE ONLY: nothing in the product includes this header.
 */
#ifndef OX_PRIVATE_H
#define OX_PRIVATE_H
#include <stdint.h>

#include "ibx_oracle.h"

#define MAXNAB 1000      /* main.F90:156 */
#define NONE (-1)


typedef struct {
    int ndat;        /* last data row index */
    double bin;
    double *x, *V, *F; /* ndat+2 rows: 0..ndat data, ndat+1 zero row */
} tab_t;

typedef struct { int n; int *data; int cap; } ivec;

typedef struct {
    int n;             /* members */
    int *idx;          /* global index of member A */
    double rcut, rlist;
    int exclusion;
    int *nneb;         /* per member */
    int *off;          /* per member: start of its list in nb */
    int *nb;           /* neighbour global indices */
    int64_t nnb;
} group_t;

/* per-owner CSR of bonded terms */
typedef struct { int *off; int *a, *b, *c; int n; } terms_t;

struct ox_ctx {
    int natoms, nvs, nitems, ntype, hybrid;
    double *mass, *invmass;
    int *tlabel;                   /* per type 1/2 */
    int *itype, *type_label;       /* per item */
    int *nbonds, *jbond;
    double *s, *r, *v, *vt, *f, *fnb;
    double box[3], boxinv[3], box2[3];
    int nbt, nat, ntt, nnbt, nopt;
    tab_t *bt, *at, *tt, *nt, *ot;
    int *ibondt, *iangt, *itort, *inbondt, *ioopt;
    int *bond_lab, *angle_lab, *tors_lab, *oop_lab, *nb_lab;
    terms_t ang, noang, tor, ftor, oop;
    int *conn_off, *conn;
    group_t g[2];
    int ngroup;
    int numatoms, numbeads, ncoarse;
    int *atom_list, *bead_list;
    /* virtual sites */
    int *vs_type, *vs_center, *vs_off, *vs_atoms;
    double *vs_invmass, *vs_coeff;   /* per VS type; coeff[vtype*ntype + atype] */
    int *vtype_natoms;
    /* control */
    int ensemble, nupdate, halt_drift, nmts, mts_quirk;
    double pressure, pfac;      /* NPT barostat: target pressure, BETA*DT*LIMAVP/TAUP (reduced) */
    double sp[10];              /* the last LIMAVP instantaneous pressures (SCALEBP.f90:18-21) */
    int iavp;                   /* SCALEBP.f90's IAVP: entries written so far, modulo LIMAVP */
    double dt, temp_in, fac;
    double mktemp, mktemp_atom, mktemp_bead;
    /* observables */
    double vnb[3], vbond[3], vang[3], vtor[3], voop[3];
    double pt[6];   /* 11 22 33 12 13 23 */
    double ek[2], temp, temp_atom, temp_bead;
    long overext;
    /* multiple time step (r-RESPA, no counterpart in the reference): nonbonded forces of the last
     * evaluation, its energy and virial */
    int respa;
    double *respa_f;         /* [nitems][3] */
    double respa_vnb[3], respa_pt[6];
    /* MTS history */
    double *mts_f;           /* [3][ncoarse][3] */
    double mts_vnb[3][3], mts_pt[3][6];
    char err[256];
};

#endif
