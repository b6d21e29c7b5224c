/*
 * FP64 CPU restatement of IBIsCO's per-step force and integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle for the HIP
 * implementation in ibisco_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it; the product never links or calls it.
 *
 * Every routine restates one reference subroutine (file:line given at each
 * function in ibx_oracle.c) in double precision with 0-based indices.  The
 * reference itself is single precision (module_var.F90:14-20,102-104).
 */
#ifndef IBX_ORACLE_H
#define IBX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Raw tables: concatenated rows of (x, V) in reduced energy units, as read from
 * the table files (before the load-time shift of RDINTERACT.f90).  Table t owns
 * rows [off[t], off[t]+nrow[t]). types[t*k ..] are 0-based particle types. */
typedef struct {
    int ntab;
    const int *types;     /* ntab * k */
    const int *off;
    const int *nrow;
    const double *x;
    const double *v;
} ox_tables;

typedef struct {
    /* sizes */
    int natoms, nvs, ntype;
    /* per type */
    const double *mass;        /* reduced */
    const int *label;          /* 1 = atom ('A'), 2 = bead ('B') */
    /* per atom */
    const int *itype;          /* natoms, 0-based */
    const int *nbonds;         /* natoms */
    const int *jbond;          /* natoms*4, 0-based, file order */
    const double *s;           /* natoms*3 unwrapped positions */
    const double *v;           /* natoms*3 reduced velocities */
    double box[3];
    /* tables */
    ox_tables bond, angle, tors, nb, oop;
    /* virtual sites: type, centre atom (-1 = centre of mass), members in CSR */
    const int *vs_type;
    const int *vs_center;
    const int *vs_off;         /* nvs+1 */
    const int *vs_atoms;
    int nvirtbonds;
    const int *virtbonds;      /* nvirtbonds*2, global 0-based */
    int nvirtangles;
    const int *virtangles;     /* nvirtangles*3 */
    /* control */
    int ensemble;              /* 0 NVE, 1 NVT, 2 NPT */
    double dt, temp_in, fac;   /* reduced dt, target T, dt/taut */
    double rcut_atom, rlist_atom, rcut_bead, rlist_bead;
    int nonbexc_atom, nonbexc_bead;
    int nupdate, halt_drift, nmts;
    int mts_quirk;             /* 1 = reproduce MOD==0 backward extrapolation */
    double pressure, pfac;     /* NPT: target pressure and BETA*DT*LIMAVP/TAUP, reduced */
} ox_desc;

typedef struct ox_ctx ox_ctx;

ox_ctx *ox_create(const ox_desc *d, char *err, int errlen);
void ox_destroy(ox_ctx *c);

/* Builds the Verlet lists of both groups (UPDATE_NEIGHBOURLIST).  mode 0 = the
 * reference's 13-cell linked-cell build, mode 1 = O(N^2) brute force. */
int ox_update_lists(ox_ctx *c, int mode);
/* Canonical pair set of the current lists: unordered pairs (lo<hi) sorted,
 * group 0 = atoms, 1 = beads+VS.  Returns the count; writes up to cap pairs. */
int64_t ox_pairs(ox_ctx *c, int group, int *out, int64_t cap);
/* Per-particle exclusion lists (CONNECTED_TO): CSR over nitems. */
int ox_exclusions(ox_ctx *c, int *off, int *idx, int cap);
/* Enumerated bonded terms: kind 0 angles (i,j,k), 1 torsions (i,j,k,l),
 * 2 parameterless angles, 3 parameterless torsions, 4 OOP. */
int ox_terms(ox_ctx *c, int kind, int *out, int cap);

void ox_shift(ox_ctx *c);
void ox_virtual_def(ox_ctx *c);
void ox_new_force(ox_ctx *c);
void ox_move(ox_ctx *c);
void ox_momentum(ox_ctx *c);
/* Multiple time step of non-hybrid systems: nonbonded forces on steps 1, k+1, ... with weight k, bonded
 * terms every step (r-RESPA; no counterpart in the reference).  Returns -1 for k < 1 or hybrid systems. */
int ox_set_respa(ox_ctx *c, int k);
/* nsteps of NEW_LOOP (or MTS_LOOP when nmts>0) starting after step0. */
int ox_run(ox_ctx *c, int nsteps, int step0);

/* Accessors.  obs layout (double[32]):
 *   0..2 V_NB, 3..5 V_BOND, 6..8 V_ANGLE, 9..11 V_TORSION, 12..14 V_OOP,
 *   15..20 PT11 PT22 PT33 PT12 PT13 PT23, 21..22 EK, 23 TEMP, 24 TEMP_ATOM,
 *   25 TEMP_BEAD, 26 volume, 27 bond overextension count, 28..30 the box */
void ox_get_obs(ox_ctx *c, double *obs);
int ox_nitems(ox_ctx *c);
void ox_get_forces(ox_ctx *c, double *f);   /* nitems*3 */
void ox_get_state(ox_ctx *c, double *s, double *v, double *r); /* s,r: nitems*3, v: natoms*3 */
void ox_set_state(ox_ctx *c, const double *s, const double *v);
/* Expanded force tables (after FTABLE): kind 0 bond, 1 angle, 2 tors, 3 nb, 4 oop.
 * Writes rows x, V, F for table t; returns the row count (NDAT+2 incl. zero row). */
int ox_table(ox_ctx *c, int kind, int t, double *x, double *V, double *F, int cap);

#ifdef __cplusplus
}
#endif
#endif
