// Device data layout shared by the kernels and the host orchestration.
//
// A context holds S systems.  Items (atoms then virtual sites) of system s
// occupy the padded range [item_off, item_off + nitems_pad); each nonbonded
// group g keeps a cell-sorted copy of its members in the padded slot range
// [slot_off[g], slot_off[g] + nslot_pad[g]).  Ranges are padded to the block
// size so that every workgroup belongs to exactly one system.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ibx_index.h"

namespace ibx {

constexpr int kBlock = 256;        // threads per workgroup for every kernel
constexpr int kWave = 64;
constexpr int kXcd = 8;            // accelerator complex dies of an MI355X, each with its own L2
constexpr int kLdsRows = 3072;     // float4 rows of nonbonded table staged in LDS (48 KiB)
constexpr int kNbPart = 16;        // doubles per workgroup partial (nonbonded)
constexpr int kBdPart = 24;        // doubles per workgroup partial (bonded: energies, overextension, virial)
constexpr int kMvPart = 16;        // doubles per workgroup partial (integrate: 15 fields)
constexpr int kObs = 32;
constexpr int kFields = 43;        // reduced fields per system: 9 nonbonded, 19 bonded, 15 integrate
constexpr int kMaxNbTabs = 64;     // nonbonded tables per system
constexpr int kMaxNbTypes = 32;    // particle types per system with several nonbonded tables
enum TermKind { KBOND = 0, KANGLE = 1, KTORS = 2, KOOP = 3 };
constexpr int kLanes = 4;                 // lanes per atom in the nonbonded kernel
constexpr int kAtomsPerWave = kWave / kLanes;
constexpr int kNbAtomsPerBlock = kBlock / kLanes;
constexpr int kListLanes = 16;            // lanes testing the candidates of one member in the list build
constexpr int kListAtoms = kBlock / kListLanes;
constexpr int kCellDiv = 3;               // cells per rlist along x and y (cell edge >= rlist/3)
constexpr int kCellDivZ = 6;              // cells per rlist along z: a column's candidate window ends closer to rlist
constexpr int kReach = kCellDiv;          // the list stencil: +-kReach columns around the member's cell in x and y,
constexpr int kReachZ = kCellDivZ;        // and +-kReachZ cells of each column in z
constexpr int kStencil = 2 * kReach + 1;
constexpr int kListRanges = 2 * kStencil * kStencil;   // candidate slot ranges per member (columns, at most 2 each)
// cells per rlist along axis k, and the fewest cells along it for the stencil (fewer: every member scans the group)
__host__ __device__ constexpr int cell_div(int k) { return k == 2 ? kCellDivZ : kCellDiv; }
__host__ __device__ constexpr int min_cells(int k) { return 2 * cell_div(k) + 1; }
constexpr double kInnerFrac = 1.0 / 3.0;  // delta = this fraction of the list skin rlist - rc
// NPT (SCALEBP.f90): LIMAVP pressures averaged per box rescaling (ALLOCATEVAR.f90:17); the per-system state holds
// that ring, then the scale factor of the step just finished and whether it rescales the box
constexpr int kLimAvp = 10;
constexpr int kNptState = 16;
constexpr int kNptMiu = kLimAvp, kNptGo = kLimAvp + 1;
constexpr int kQueueInts = kXcd + 1;       // work queues of the persistent force launch, then its finished count
constexpr int kTileCandidates = 4;         // force launch layouts the tile autotuner times

// One bonded contribution slot: the force of one term on one member (12 bytes; the owning item gathers its slots).
struct Contrib {
    float x, y, z;
};

// Position of entry k of sorted slot i in the Verlet list.  Entries of the
// kAtomsPerWave slots of one wave are interleaved so that the kLanes lanes
// serving a slot, and the slots of a wave, read consecutive words.
// A slot's cap entries hold two runs: nneb.x pairs that were closer than
// rc + delta when the list was built, from k = 0 up, and nneb.y pairs with
// rc + delta <= r < rlist, from k = cap - 1 down.
__host__ __device__ __forceinline__ size_t list_pos(int slot, int k, int cap) {
    return (size_t)(slot / kAtomsPerWave) * kAtomsPerWave * cap + (size_t)(k / kLanes) * kWave +
           (slot % kAtomsPerWave) * kLanes + (k % kLanes);
}

struct DevGroup {
    int slot_off, nslot_pad, nmem;
    int mem_off;                   // offset into the member list
    int cell_off, ncell;           // global cell range
    int nc[3];
    int ncmax[3];                  // the grid the cell range was sized for (NPT: the grid never grows past it)
    int all_pairs;                 // fewer than min_cells(k) cells along an axis: every member scans the group
    double rl2;                    // rlist^2 (FP64 list criterion)
    float rl2u_lo, rl2u_hi;        // FP32 pre-classification thresholds, working units^2
    float rc2u;                    // rcut^2 in working units^2
    float rc2u_lo, rc2u_hi;        // band around rc2u where the FP64 cutoff test decides
    double rc2;                    // rcut^2 (FP64 cutoff criterion)
    float rin2u;                   // (rc + delta)^2: pairs closer than this at the rebuild form the front run
    float dmax2u;                  // (delta/2)^2 less a margin: while every member moved less, the back run
                                   // cannot reach rc and is skipped
    int wg_off, nwg;               // workgroup range of this system in the group's grid
};

struct DevSys {
    double box[3], boxinv[3], box2[3];
    double inv_box_frac[3];        // 1/L per dimension for fixed-point conversion
    double unit;                   // working length unit: L_x / 2^32 (nm)
    float rho[3];                  // L_k / L_x (1 for a cubic box)
    int item_off, nitems_pad, natoms, nitems, nvs;
    int item_wg_off, item_nwg;
    int term_wg_off, term_nwg;     // workgroups of this system's bonded terms
    int hybrid;
    DevGroup g[2];
    int ngroup;
    // nonbonded tables of this system
    int nb_row_off, nb_nrow;       // rows in the global float4 table array
    int nb_ntab, nb_tab_off;       // table headers
    int ntype, nbmap_off;          // ntype*ntype map -> global table header index
    // integration
    double dt, temp_in, fac;
    int ensemble;                  // 0 NVE, 1 NVT, 2 NPT
    double pressure, pfac;         // NPT: target pressure and BETA*DT*LIMAVP/TAUP, reduced units
    double mktemp, mktemp_atom, mktemp_bead;
    double volume;
    int halt_drift, nmts, mts_quirk, nupdate;
    int vs_off;                    // offset into the VS arrays
    int bead_off;                  // offset into MTS bead buffers
    int nbead;                     // members of group 1 (beads+VS) for MTS
    int live;                      // 0 once retired: the system's workgroups return at once and its
                                   // state, observables, window sums and RDF keep their values
    int step0;                     // context step at which this system's own step 0 was (a system assigned
                                   // to a slot of a running batch); MOMENTUM counts the system's own steps
    int last_step;                 // the system's own last step: later steps leave it as after that step
    const int *step_ctr;           // the context's step counter (steps completed; k_move advances it)
};

// Flag on the workgroup -> system words of the integrator and bonded launches:
// the system is retired, so the workgroup returns before loading anything.
constexpr int kWgRetired = 1 << 30;

// Whether the system takes part in step t of the context (t = the counter + 1
// for work before the move of step t).
__host__ __device__ inline bool alive_at(const DevSys &S, int t) { return S.live && t - S.step0 <= S.last_step; }

struct NbTab {                     // one nonbonded table as seen by a system
    int row_off, ndat;
    double c;                      // unit / bin: table coordinate per working length unit
    float bin;
    int uniform;                   // abscissas are k*bin (else nb_xoff holds x_k/bin - k per row)
};

struct BdTab {                     // one bonded table (bond/angle/torsion/OOP)
    int row_off, ndat;
    double x0, bin, inv_bin;
    int label;                     // energy class 1..3
    int pad;
};

// Fixed-point fractional coordinate: floor(frac(s/L) * 2^32).
__device__ __forceinline__ uint32_t to_fixed(double s, double inv_l) {
    double t = s * inv_l;
    t = t - floor(t);
    double u = t * 4294967296.0;
    if (u >= 4294967296.0) u = 0.0;
    return (uint32_t)u;
}

}  // namespace ibx
