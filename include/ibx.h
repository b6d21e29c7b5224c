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
#define IBX_E_PENDING (-7)       /* ibx_server_wait: the job has not finished within the timeout */

#define IBX_NTILES 3              /* ibx_set_tiles: force launch layout, blocks per take, workgroups per CU */
#define IBX_TILE_CANDIDATES 4     /* force launch layouts the tile autotuner times */

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
    int ensemble;              /* 0 NVE, 1 NVT (Berendsen), 2 NPT (Berendsen thermostat and barostat) */
    double dt, temp_in, fac;   /* reduced time step, target temperature, dt/taut */
    double rcut_atom, rlist_atom, rcut_bead, rlist_bead;
    int nonbexc_atom, nonbexc_bead; /* 4: exclude 1-2,1-3; 5: also 1-4 */
    int nupdate, halt_drift, nmts;
    int mts_quirk;             /* 1 reproduces MTS_LOOP.f90:100 at MOD==0 */
    /* ensemble 2 (NPT, Berendsen barostat SCALEBP.f90): target pressure and
     * PFAC = BETA*DT*LIMAVP/TAUP, reduced (RDCONTROL.f90:107-108,317-337,
     * ALLOCATEVAR.f90:17) */
    double pressure, pfac;
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
 * of a batch advance together.  A hybrid system with nmts > 0 runs the MTS
 * cycle of MTS_LOOP.f90 instead of NEW_FORCE: the bead group's nonbonded forces,
 * energies and virial are evaluated on steps with MOD(step, 3+nmts) = 1, 2, 3
 * and extrapolated on the others, with I = 0 on the last step of a cycle as in
 * the reference when mts_quirk is set (else I = 3+nmts). */
int ibx_step(ibx_ctx *ctx, int nsteps, int step0);

/* The same steps as ibx_step, bit for bit, launched as captured HIP graphs:
 * blocks of four list intervals (and multiple-time-step cycles), then of one,
 * then eager steps for the rest (batched replicas and serving use this path).
 * A block may start at any step: graphs are kept per block length and phase
 * of step0 in the block period (the period is ibx_get_counters()[4]), so a
 * caller stepping in chunks of whole periods replays after the first call. */
int ibx_step_graph(ibx_ctx *ctx, int nsteps, int step0);

/* The steps of ibx_step, bit for bit, for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The reduction partials and field sums the
 * kernels write (per-workgroup partial sums of every kernel class and the
 * per-system field sums) live in a static 16 MiB device array of
 * the library instead of the context's buffers; the call copies the context's
 * into the array first and back at the end, and allocates nothing.  A context
 * needing more than 16 MiB of them (about 2 million particles) is rejected
 * with IBX_E_INVALID.  The calling thread's current device must be the
 * context's.  Calls hold a process-wide lock and wait for their stream, so two
 * calls never overlap on the array (a call that overlapped another on the same
 * GPU would race on it). */
int ibx_step_noalloc(ibx_ctx *ctx, int nsteps, int step0);

/* Step counters of the context since creation: [0] steps run eagerly, [1]
 * steps replayed from graphs, [2] graph launches, [3] graph captures; [4] the
 * block period (steps), [5] steps of a long block, [6..7] 0.  No counterpart
 * in the reference (instrumentation). */
int ibx_get_counters(ibx_ctx *ctx, int64_t out[8]);

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
/* The current box edges of system sys (nm): an NPT system's box is rescaled on
 * the device every LIMAVP = 10 steps (SCALEBP.f90:36-40; the reference prints
 * it with its averages, AVERAGE.f90:23-26).  obs[26] is its volume. */
int ibx_get_box(ibx_ctx *ctx, int sys, double box[3]);
/* Window sums of obs[0..25] over the steps since the last reset (AVERAGE.f90). */
int ibx_get_averages(ibx_ctx *ctx, int sys, double sums[32], int reset);

/* Per-item forces (natoms+nvs)*3 after the last force phase, reduced units. */
int ibx_get_forces(ibx_ctx *ctx, int sys, double *f);
/* State of system sys, xyz interleaved: s (natoms+nvs)*3 unwrapped positions
 * (SXYZ), v natoms*3 reduced velocities (VXYZ).  Either pointer may be NULL. */
int ibx_get_state(ibx_ctx *ctx, int sys, double *s, double *v);
/* Replaces the state of system sys (same layout; virtual sites are placed from
 * their atoms).  New positions rebuild the lists of every system of the
 * context at their current positions (main.F90:234 after RDCOOR). */
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

/* Sizes: info[0] natoms, [1] nvs, [2]/[3] list capacity of groups 0/1,
 * [4] padded slots of group 0, [5] device bytes the context holds,
 * [6] systems in the context, [7] padded items of all systems. */
int ibx_get_info(ibx_ctx *ctx, int sys, int64_t info[8]);

/* Multiple time step (r-RESPA in leap-frog form; no counterpart in the
 * reference, whose MTS_LOOP.f90 extrapolates the CG forces of hybrid models):
 * every system of the context evaluates its nonbonded forces on steps 1,
 * k+1, 2k+1, ... and applies them with weight k (an impulse of k inner
 * steps), and its bonded terms on every step; dt stays the inner step.
 * Observables on the other steps carry the last nonbonded energy and virial.
 * k = 1 (the default) is the plain step loop, bit for bit. */
int ibx_set_respa(ibx_ctx *ctx, int k);

/* List-skin autotuner (on by default; no counterpart in the reference).  At
 * every scheduled rebuild each system's inner skin delta -- pairs closer than
 * rc + delta at the build are read every step, the rest of the list (to
 * rlist) only once a member has moved 0.49 delta since the build -- is chosen
 * from the largest displacement that system reached over the interval just
 * ended, to minimise the list entries read over the next one.  The listed
 * pairs and the forces do not depend on delta beyond FP32 summation order;
 * the choice depends only on the system's own trajectory, so batches stay
 * bit-identical to their systems run alone.  Off: delta = (rlist - rc)/3
 * from the next scheduled rebuild on (delta changes only where a list is built). */
int ibx_set_skin_tuning(ibx_ctx *ctx, int on);
/* Current inner skin delta (nm) of each group of system sys. */
int ibx_get_skin(ibx_ctx *ctx, int sys, double delta[2]);

/* Tile autotuner and the layout of the force launch (no counterpart in the
 * reference).  The force phase of a step (nonbonded blocks of 64 slots, bonded
 * blocks of 256 terms, the previous step's field sums) runs either as one
 * workgroup per block (tiles[0] = 0) or as a persistent grid of tiles[2]
 * workgroups per CU that take tiles[1] blocks at a time from per-XCD device
 * work queues (tiles[0] = 1), keeping a system's nonbonded tables in LDS from
 * one block to the next.  Every layout computes each block the same way, so
 * results are the same bits.  tiles[0] = -1 (the default) leaves the choice to
 * the autotuner: at the first ibx_step / ibx_step_graph call after creation
 * (or after this call) it times one-workgroup-per-block and persistent grids
 * of 5 or 4 workgroups per CU taking 1 or 2 blocks, on the context's own
 * systems, and keeps the fastest.  Values -1 in tiles[1] and tiles[2] take 1
 * and 5. */
int ibx_set_tiles(ibx_ctx *ctx, const int tiles[IBX_NTILES]);
/* The layout in use (all -1 until the first step call applies or tunes it)
 * and, when ms is not null, the autotuner's time per force launch of each
 * candidate in milliseconds (-1: not timed): one workgroup per block,
 * persistent 5 per CU taking 1 block, 5 per CU taking 2, 4 per CU taking 1. */
int ibx_get_tiles(ibx_ctx *ctx, int tiles[IBX_NTILES], float ms[IBX_TILE_CANDIDATES]);

/* Stops advancing system sys of a batch: its workgroups return at once, and
 * its state, observables, window sums and RDF keep the values of its last step
 * until ibx_assign gives the place to another system.  The other systems are
 * unaffected (bit for bit).  No counterpart in the reference (one system per
 * run); the serving loop uses it when a job of a batch reaches its step count. */
int ibx_retire(ibx_ctx *ctx, int sys);

/* System sys of a batch advances through its own step nsteps (counted from the
 * step0 it started at: 0, or that of ibx_assign) and then stops by itself: its
 * workgroups return at once and its state, observables, window sums and RDF
 * keep the values of that step while the context steps on.  nsteps < 0: no
 * limit (the default).  The serving loop sets it so that a batch always steps
 * in whole graph blocks and collects a job's results at the next block
 * boundary.  The reference runs one system for NSTEP steps (RDCONTROL); this is
 * that limit per member of a batch. */
int ibx_set_last_step(ibx_ctx *ctx, int sys, int nsteps);

/* Gives place sys of a batch, whose system has been retired, to the system of
 * desc from context step step0 on (the step0 the next ibx_step* call passes; a
 * multiple of update_neighbour_list and of the RESPA factor).  The system must
 * fit what the place holds (ibx_get_capacity; a place holds its first system's
 * sizes, or the reservation its creator asked for), have no virtual sites and
 * one nonbonded group, and share the batch's update_neighbour_list.  Its own
 * steps count from step0 (MOMENTUM, averages), its observables, window sums
 * and RDF start at zero, and its first lists are built at once; the other
 * systems of the batch continue bit for bit, and captured graphs stay valid.
 * IBX_E_NOMEM when it does not fit, IBX_E_INVALID for the other conditions.
 * No counterpart in the reference (one system per run): the serving loop uses
 * it to admit a queued job into a running batch. */
int ibx_assign(ibx_ctx *ctx, int sys, const ibx_system_desc *desc, int step0);

/* What place sys of a batch holds: [0] items (padding included), [1] bonded
 * terms, [2] term member slots, [3] exclusion entries, [4..5] members of
 * nonbonded groups 0 and 1, [6..7] list cells of each group, [8] nonbonded
 * table rows, [9] nonbonded tables, [10] type-pair map entries, [11] bonded
 * tables, [12] bonded table rows; [13] the batch's list capacity per member,
 * [14] steps of a graph block (ibx_assign's step0 alignment in the serving
 * loop). */
int ibx_get_capacity(ibx_ctx *ctx, int sys, int64_t cap[16]);

/* ---- Trajectory and restart output (WRITETRJ.F90) ------------------------
 * Host-side writers of the reference's output formats, fed from the device
 * state at output steps only (the window sums stay on the device between
 * them).  Errors are reported through ibx_output_last_error. */

/* Appends one TRZ frame to path as WRITETRJ.F90:31-104 writes it (frame 1
 * creates the file and writes the header records: the title in 80
 * characters, then 10).  Fortran unformatted sequential records, each between
 * 4-byte length markers: frame, frame*ntrj, natoms, time (ps); the box as a
 * 3x3 matrix (nm); pressure and tensor in kPa (p, 11, 12, 22, 13, 23, 33); 6,
 * then total, potential and kinetic energy (kJ/mol), the temperature as
 * SUM(EK)*MKTEMP (reduced, as the reference writes it), 0, 0; x, y, z (nm) and
 * vx, vy, vz (nm/ps) in single precision.  Inputs in reduced units: pt the
 * pressure tensor (11 22 33 12 13 23), epot and ek the potential and kinetic
 * energy, mktemp as in ibx_system; s natoms*3 (nm) and v natoms*3 (reduced),
 * xyz interleaved. */
int ibx_trz_write_frame(const char *path, const char *title, int frame, int ntrj, int natoms, double time_ps,
                        const double box[3], const double pt[6], double epot, double ek, double mktemp,
                        const double *s, const double *v);
/* The restart file of WRITETRJ.F90:115-178, which RDCOOR.f90 reads back as a
 * coordinate file: title, time (ps), box, the record description, then per
 * molecule (nmol of natm[m] atoms named names[m]) each atom's index, type,
 * bond count and position (nm), and its velocity (nm/ps) with its bonded
 * atoms.  itype (0-based), nbonds and jbond (natoms*4, 0-based) as in
 * ibx_system_desc; v reduced. */
int ibx_restart_write(const char *path, const char *title, double time_ps, const double box[3], int natoms,
                      const int *itype, const int *nbonds, const int *jbond, const double *s, const double *v,
                      int nmol, const int *natm, const char *const *names);
/* A TRZ frame of system sys from its current state and observables
 * (ibx_get_state, ibx_get_observables after a step); time = step * dt +
 * initime_ps (WRITETRJ.F90:57). */
int ibx_write_trz(ibx_ctx *ctx, int sys, const char *path, const char *title, int frame, int ntrj, int step,
                  double initime_ps);
/* The restart file of system sys from its current state (molecules as in
 * ibx_restart_write). */
int ibx_write_restart(ibx_ctx *ctx, int sys, const char *path, const char *title, int step, double initime_ps,
                      int nmol, const int *natm, const char *const *names);
const char *ibx_output_last_error(void);

/* ---- Serving loop (no counterpart in the reference) ----------------------
 * Jobs (one system and a step count each) are queued in arrival order.  A
 * queued job goes to a free place of a running batch that holds it (the
 * smallest such place), where it starts at the batch's next graph-block
 * boundary (ibx_assign); otherwise, while the estimated device bytes of all
 * batches stay within hbm_budget and fewer than max_cohorts batches run, a new
 * batch is created from the queue.  Each place of a batch is reserved for a
 * size class (1024 items times a power of class_ratio, the other sizes scaled
 * from its job), and a new batch also keeps spare_places free places per size
 * class seen so far, for jobs that arrive while it runs.  Batches run
 * concurrently on their own streams and host threads.  A job that reaches its
 * step count is retired, its results are copied out and its place is free for
 * the next job; a batch with no job left ends and its bytes return to the
 * budget.  When the oldest queued job fits no free place and no new batch can
 * start, the batch with the least work left stops taking jobs, so that it ends
 * and makes room.  Each job's results equal those of the same system stepped
 * alone in its own context (ibx_step_graph from step 0, RDF every rebuild). */
typedef struct ibx_server ibx_server;

typedef struct {
    int64_t hbm_budget;   /* device bytes for all resident jobs (<= 0: 200e9; capped at 90% of free memory) */
    int max_cohorts;      /* batches stepping at once (<= 0: 3) */
    int max_cohort_jobs;  /* jobs per batch (<= 0: 256) */
    int chunk_steps;      /* most steps between completion checks of a batch (<= 0: 200) */
    int rdf_nbin;         /* RDF bins per job (0: none) */
    double rdf_dr;        /* RDF bin width, nm */
    double step_band;     /* > 1: a new batch takes, besides the oldest queued job, only jobs whose step counts
                             are within this factor of its own, so its members end close together (<= 1: off) */
    int spare_places;     /* free places per size class seen so far in a new batch (0: 1); < 0: places
                             are not handed over (a batch takes jobs only when it is created) */
    double class_ratio;   /* growth between size classes of places (<= 1: 1.25) */
} ibx_server_config;

typedef struct {
    int status;           /* IBX_OK or the error of the job's batch */
    int natoms, nsteps;
    int cohort, cohort_jobs;      /* batch that ran the job and its number of jobs */
    double t_submit, t_admit, t_done;   /* seconds since ibx_server_create */
    double mean_pressure;         /* mean over the job's steps of (PT11+PT22+PT33)/3, reduced units */
    double obs[32];               /* as ibx_get_observables after the last step */
    double sums[32];              /* as ibx_get_averages over the job's steps */
    char error[256];
} ibx_job_result;

/* Device bytes ibx_create allocates for this one system, plus an RDF of
 * rdf_nbin bins: the footprint the serving loop admits by.  Errors are
 * reported through ibx_server_last_error. */
int ibx_estimate_bytes(const ibx_system_desc *desc, int rdf_nbin, int64_t *bytes);

/* cfg may be NULL (defaults).  Starts the scheduler thread. */
int ibx_server_create(const ibx_server_config *cfg, ibx_server **out);
/* Queues a job; the descriptor's arrays are copied before the call returns.
 * keep_state != 0 keeps the final state for ibx_server_wait. */
int ibx_server_submit(ibx_server *srv, const ibx_system_desc *desc, int nsteps, int keep_state, int64_t *job);
/* Waits up to timeout_s seconds (< 0: no limit) for a job, then hands over its
 * result once: res, rdf (rdf_nbin values of g(r)), and for keep_state jobs s
 * ((natoms+nvs)*3) and v (natoms*3) as ibx_get_state; any pointer may be NULL.
 * Returns IBX_E_PENDING while the job runs, else the job's status. */
int ibx_server_wait(ibx_server *srv, int64_t job, double timeout_s, ibx_job_result *res, double *rdf, double *s,
                    double *v);
/* [0] submitted, [1] completed, [2] failed, [3] queued, [4] running batches,
 * [5] batches created, [6] peak running batches, [7] bytes in use (estimated),
 * [8] peak bytes, [9] atom-steps of completed jobs, [10] seconds since
 * creation, [11] most places of a batch, [12] mean jobs per batch (those it
 * started with and those given its free places), [13] effective budget,
 * [14] jobs given a free place of a running batch, [15] atom-steps run so far
 * (every live place of every batch, jobs still running included). */
int ibx_server_stats(ibx_server *srv, double st[16]);
/* Host seconds the batches' threads spent, summed over batches: [0] in
 * ibx_step_graph, [1] handing places over (ibx_assign), [2] copying results
 * out, [3] creating contexts; [4] step calls, [5] hand-overs, [6] steps run
 * (each batch's steps counted once). */
int ibx_server_timing(ibx_server *srv, double t[8]);
/* Runs every submitted job to its end, then stops the threads and frees. */
void ibx_server_destroy(ibx_server *srv);
/* Message of the last failed ibx_server_* call on the calling thread. */
const char *ibx_server_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
