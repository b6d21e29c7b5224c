// Launch wrappers of kernels.hip (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "ibx_kernels.h"

namespace ibx {

struct RebuildArgs {
    const DevSys *sys;
    const int *wg_sys;
    const int *cell_sys;
    int g, nwg, ncell_total, nitems_total, cap;
    bool cubic;
    const int *mem;
    const double *s;
    const int *itype, *label;
    int *mem_cell, *mem_rank, *cell_count, *cell_start, *tmp, *perm, *slot_cell;
    int *tile_sum;      // one per kBlock cells (the cell-start scan)
    double4 *r64;
    int4 *pos4, *pos4_ref;
    int *inv_perm;
    int *dmax;          // this group's per-system largest displacement^2 since the rebuild (float bits)
    int nsys;
    // list-skin tuning in the same launch as the cell count (-1: none; 0: restore the fixed inner
    // radius; 1: tune), from every group's displacement over nmove moves
    int tune, nmove;
    DevSys *tune_sys;
    const int *dmax_all;
    const int *excl_off, *excl;
    int *list;
    int2 *nneb;
    int *flags;
    int only;           // >= 0: rebuild only this system (the others keep their lists), else every system
    // RDF of the listed group-0 pairs, accumulated by the list build (null: none)
    unsigned long long *rdf_hist;
    int *rdf_samples;
    double *rdf_ivol;   // per system: the sum over samples of 1 / volume
    int rdf_nbin;
    float rdf_inv_dr;   // 1 / bin width (nm^-1)
};

struct NbArgs {
    const DevSys *sys;
    const int *wg_sys;
    int g, nwg, cap, nitems_total, max_rows_lds;
    bool multi, cubic;
    const int *perm;
    const int4 *pos4;
    const int *list;
    const int2 *nneb;
    const int *dmax;
    int nsys;
    const float4 *rows;
    const float *xoff;
    const NbTab *tabs;
    const double *s;
    const int *map;
    float *fnb;
    double *part;
};

struct TermLaunch {
    const DevSys *sys;
    const int4 *mem;
    const int *meta;
    const int4 *slot;
    const BdTab *tabs;
    const float4 *rows;
    const double *xs;
    const double *vs_com;   // per virtual site: the mass-weighted centre of its atoms
};

struct MoveLaunch {
    double *s, *v;
    const float *fnb;
    const Contrib *contrib;
    const int *inc_off;
    int ncontrib;
    float *ftot;
    const int *itype, *label;
    const double *mass, *invmass;
    const int *vs_of;
    const double *vs_coeff;
    int4 *pos4[2];
    const int4 *pos4_ref[2];
    const int *inv_perm[2];
    int *dmax;          // [group][system] largest displacement^2 since the rebuild (float bits)
    int nsys;
    const double *sysstate;
    double *part;
    int *step;
    int nitems_total, mode;
    float nb_weight;    // weight of the nonbonded force in this move (multiple time step)
};

// The field sums and observables of one step (k_field_sums).
struct FinLaunch {
    const DevSys *sys;
    const double *part_nb0, *part_nb1, *part_bd, *part_mv;
    int snb0, snb1, sbd, smv;   // workgroups (field strides) of each partial buffer
    double *fsum;
    double *obs, *avg, *sysstate;
    const int *step;
    int mode, nsys;
    double *npt;        // [system][kNptState] barostat state
};

void launch_mts(hipStream_t st, bool save, float *fnb, const int *label, int n, double *part, int npart,
                float *mts_f, double *mts_part, int m, double I);
void launch_rebuild(hipStream_t st, const RebuildArgs &a);
void launch_nonbonded(hipStream_t st, const NbArgs &a);
void launch_terms(hipStream_t st, int nwg, const int *wg_sys, const TermLaunch &b, const double *s, int nitems_total,
                  int nterms, Contrib *contrib, int ncontrib, double *part, int *flags);
void launch_move(hipStream_t st, int nwg, const DevSys *sys, const int *wg_sys, const MoveLaunch &m);
void launch_vs_place(hipStream_t st, int nwg, const DevSys *sys, const int *wg_sys, double *s, int nitems_total,
                     const int *label, const int *itype, const int *vs_center, const int *vs_moff, const int *vs_members,
                     const double *mass, const double *vs_invmass, double *vs_com, int4 *pos4, const int4 *pos4_ref,
                     const int *inv_perm, int *dmax);
void launch_finalize(hipStream_t st, const FinLaunch &f);
// The NPT box rescaling after the field sums of a step (k_npt_positions over the item blocks, then k_npt_box).
void launch_npt(hipStream_t st, int nwg, DevSys *sys, const int *wg_sys, int nsys, double *npt, double *s,
                double *vs_com, const int *itype, const int *label, int nitems_total, NbTab *nb_tabs, int *dmax);
// The force phase of a step in one launch: the group-0 nonbonded force (nb null: none this step), the bonded
// terms, and the field sums of the previous step (fin null: none).
// How the force launch is laid out: one workgroup per block, or a persistent grid of `grid` workgroups taking
// `chunk` blocks at a time from per-XCD work queues (queue: kXcd + 1 ints, zero between launches).
struct ForceTiles {
    bool persistent = false;
    int chunk = 1;
    int grid = 0;
    int *queue = nullptr;
};
void launch_step_forces(hipStream_t st, const NbArgs *nb, int term_nwg, const int *wg_terms, const TermLaunch &b,
                        const double *s, int nitems_total, int nterms, Contrib *contrib, int ncontrib, double *part_bd,
                        int *flags, const FinLaunch *fin, const ForceTiles &ft);
// Sets n ints to v (a kernel, not a memset node, so that code a caller may capture stays replayable).
void launch_fill_int(hipStream_t st, int *p, int n, int v);


}  // namespace ibx
