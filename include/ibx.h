/*
 * ibx — MI355X (gfx950) implementation of IBIsCO's per-step force and
 * integration path behind a plain C ABI.
 *
 * One context holds one system or a batch of independent systems (replicas)
 * resident in HBM.  Indices are 0-based, lengths in nm, energies and masses in
 * IBIsCO reduced units (k_B*300 K per particle, 12 g/mol) unless stated;
 * ibx_load_dir() reads the reference's own input files and converts.
 *
 * Every call returns 0 on success or a negative IBX_E* code; the message of the
 * last failure is available from ibx_last_error().  Nothing calls exit().
 * A context is used by one host thread at a time; contexts on the same device
 * run on their own HIP streams.
 */
#ifndef IBX_H
#define IBX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IBX_OK 0
#define IBX_E_INVALID (-1)       /* invalid input (RDCONTROL/RDINTERACT/RDCOOR fatal errors) */
#define IBX_E_LIST_OVERFLOW (-2) /* "LIST TOO SMALL" (NEW_NEIGHBOUR_WITHLIST.f90:72) */
#define IBX_E_TABLE (-3)         /* torsion/OOP outside its table (TORSIONS.inc:103-109) */
#define IBX_E_DEVICE (-4)        /* HIP runtime failure, or no usable gfx950 device */
#define IBX_E_NOMEM (-5)         /* device allocation failed */
#define IBX_E_IO (-6)            /* input file missing or unreadable */

typedef struct ibx_ctx ibx_ctx;

/* Raw tables of one kind: ntab tables whose rows (x as in the table file,
 * V already in reduced units) are concatenated; table t owns rows
 * [off[t], off[t]+nrow[t]) and applies to the 0-based particle types
 * types[t*k .. t*k+k-1] (k = 2 bond/nb, 3 angle, 4 torsion/OOP). */
typedef struct {
    int ntab;
    const int *types;
    const int *off;
    const int *nrow;
    const double *x;
    const double *v;
} ibx_tables;

/* Everything the reference reads from control, interaction, coordinate and the
 * hybrid files, for one system. */
typedef struct {
    int natoms, nvs, ntype;
    const double *mass;        /* per type, reduced */
    const int *label;          /* per type: 1 atom ('A'), 2 bead ('B') */
    const int *itype;          /* per atom */
    const int *nbonds;         /* per atom */
    const int *jbond;          /* per atom, 4 slots, file order */
    const double *s;           /* natoms*3 unwrapped positions */
    const double *v;           /* natoms*3 velocities, reduced */
    double box[3];
    ibx_tables bond, angle, tors, nb, oop;
    const int *vs_type;        /* nvs */
    const int *vs_center;      /* nvs: centre atom, or -1 for a centre of mass */
    const int *vs_off;         /* nvs+1 CSR into vs_atoms */
    const int *vs_atoms;
    int nvirtbonds;
    const int *virtbonds;      /* nvirtbonds*2 */
    int nvirtangles;
    const int *virtangles;     /* nvirtangles*3 */
    int ensemble;              /* 0 NVE, 1 NVT (Berendsen) */
    double dt, temp_in, fac;   /* reduced time step, target temperature, dt/taut */
    double rcut_atom, rlist_atom, rcut_bead, rlist_bead;
    int nonbexc_atom, nonbexc_bead; /* 4: exclude 1-2,1-3; 5: also 1-4 */
    int nupdate, halt_drift, nmts;
    int mts_quirk;             /* 1 reproduces MTS_LOOP.f90:100 at MOD==0 */
} ibx_system_desc;

/* Replaces the set-up half of main.F90:107-234 (RDCONTROL .. UPDATE_NEIGHBOURLIST):
 * builds terms (SETLIS.f90), exclusions (BUILD_CONNECTIVITY.f90, RDVIRTBONDS.f90),
 * force tables (FTABLE.f90), uploads the state and builds the first lists.
 * nsys > 1 creates a batch of independent replicas advanced together. */
int ibx_create(int nsys, const ibx_system_desc *const *desc, ibx_ctx **out);

/* Reads a reference-format input directory (control, interaction and its
 * table files, coordinate, virtual, virtbonds, virtangles) and creates a
 * one-system context: RDCONTROL.f90, RDINTERACT.f90, RDCOOR.f90, RDVIRTUAL.f90. */
int ibx_load_dir(const char *dir, ibx_ctx **out);

void ibx_destroy(ibx_ctx *ctx);
const char *ibx_last_error(const ibx_ctx *ctx);   /* ctx may be NULL: last create error */

/* nsteps of the step loop (NEW_LOOP.F90:19-156), steps numbered
 * step0+1..step0+nsteps: SHIFT, VIRTUAL_DEF, UPDATE_NEIGHBOURLIST every
 * nupdate, NEW_FORCE, MOVE, MOMENTUM every halt_drift, AVERAGE.  All replicas
 * of a batch advance together.  The MTS cycle (MTS_LOOP.f90, nmts>0) is not on
 * the device: ibx_create rejects such systems with IBX_E_INVALID. */
int ibx_step(ibx_ctx *ctx, int nsteps, int step0);

/* The same steps as ibx_step, launched as a captured HIP graph replayed per
 * list interval (batched replicas and serving use this path). */
int ibx_step_graph(ibx_ctx *ctx, int nsteps, int step0);

/* NEW_FORCE.F90:7-83 at the current positions without moving: fold, virtual
 * sites, (re)build the lists when rebuild != 0, forces, energies, virial. */
int ibx_compute_forces(ibx_ctx *ctx, int rebuild);

/* UPDATE_NEIGHBOURLIST.F90:7-88 at the current positions. */
int ibx_rebuild_lists(ibx_ctx *ctx);

/* Observables of system sys after the last step or force call, FP64:
 *   0..2 V_NB, 3..5 V_BOND, 6..8 V_ANGLE, 9..11 V_TORSION, 12..14 V_OOP,
 *   15..20 PT11 PT22 PT33 PT12 PT13 PT23 (pressure tensor after MOVE, virial
 *   only after ibx_compute_forces), 21..22 EK, 23 TEMP, 24 TEMP_ATOM,
 *   25 TEMP_BEAD, 26 volume, 27 bond over-extension count,
 *   28 steps accumulated in the window sums. */
int ibx_get_observables(ibx_ctx *ctx, int sys, double obs[32]);
/* Window sums of obs[0..25] over the steps since the last reset (AVERAGE.f90). */
int ibx_get_averages(ibx_ctx *ctx, int sys, double sums[32], int reset);

/* Per-item forces (natoms+nvs)*3 after the last force phase, reduced units. */
int ibx_get_forces(ibx_ctx *ctx, int sys, double *f);
/* State of system sys, xyz interleaved: s (natoms+nvs)*3 unwrapped positions
 * (SXYZ), v natoms*3 reduced velocities (VXYZ).  Either pointer may be NULL. */
int ibx_get_state(ibx_ctx *ctx, int sys, double *s, double *v);
int ibx_set_state(ibx_ctx *ctx, int sys, const double *s, const double *v);
/* Replaces nonbonded table t of system sys (raw rows as in ibx_tables). */
int ibx_set_nb_table(ibx_ctx *ctx, int sys, int t, int nrow, const double *x, const double *v);
/* Switches ensemble/thermostat of every system: ensemble 0 NVE / 1 NVT. */
int ibx_set_thermostat(ibx_ctx *ctx, int ensemble, double temp_in, double fac, double dt);

/* Canonical pair set of the current Verlet list of group g (0 atoms,
 * 1 beads + virtual sites): unordered pairs (lo<hi) of local indices, sorted.
 * With out == NULL only *count is written. */
int ibx_get_pairs(ibx_ctx *ctx, int sys, int group, int *out, int64_t cap, int64_t *count);

/* Radial distribution function of each system, accumulated on the device at
 * every list rebuild while every > 0 (0 pauses it), in nbin bins of width dr
 * over the group-0 pairs of the list: nonbonded pairs only (exclusions are
 * not in the list), and bins must stay below rlist.  Systems
 * of a batch may have different boxes.  ibx_rdf_get normalises by the ideal
 * gas shell counts and the number of accumulated rebuilds. */
int ibx_rdf_enable(ibx_ctx *ctx, int nbin, double dr, int every);
int ibx_rdf_get(ibx_ctx *ctx, int sys, double *g, int reset);

/* Device-side timing of the last ibx_step call in ms per kernel class:
 * [0] nonbonded, [1] bonded, [2] integrate, [3] list rebuild, [4] whole step. */
int ibx_get_timing(ibx_ctx *ctx, double ms[8]);

/* Enables (on != 0) or disables device timing of ibx_step and clears the totals. */
int ibx_set_timing(ibx_ctx *ctx, int on);

/* Sizes: natoms, nvs, list capacity, padded slots; and device bytes held. */
int ibx_get_info(ibx_ctx *ctx, int sys, int64_t info[8]);

#ifdef __cplusplus
}
#endif
#endif
