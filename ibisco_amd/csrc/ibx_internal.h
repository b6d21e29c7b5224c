// Internal host-side data model of the ibx library (not part of the C ABI).
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "ibx.h"

namespace ibx {

// A tabulated interaction after the load-time shift of RDINTERACT.f90:450-537
// and the differentiation of FTABLE.f90:9-91.  Rows 0..ndat hold data; row
// ndat+1 is the zero row the reference reads past the end of its tables.
struct Table {
    int ndat = 0;
    double bin = 0.0;
    double x0 = 0.0;
    std::vector<double> x, V, F;
};



// Sizes of one system's share of a batch's device arrays: what a place in a
// batch holds, and what a system needs to take a place over.
struct SlotSizes {
    int items = 0;                       // items (atoms, beads, virtual sites), padding included
    int terms = 0, contrib = 0;          // bonded terms; their members' contribution slots
    int excl = 0;                        // exclusion entries
    int members[2] = {0, 0};             // nonbonded group members (list slots)
    int cells[2] = {0, 0};               // list cells of each group
    int nb_rows = 0, nb_tabs = 0, nb_map = 0;   // nonbonded table rows, headers, type-pair map entries
    int bd_tabs = 0, bd_rows = 0;        // bonded table headers and rows
};

struct HostSystem {
    int natoms = 0, nvs = 0, nitems = 0, ntype = 0;
    bool hybrid = false;
    std::vector<double> mass, invmass;
    std::vector<int> tlabel;            // per type: 1 atom, 2 bead
    std::vector<int> itype, type_label; // per item: 1 atom, 2 bead, 3 virtual site
    std::vector<int> nbonds, jbond;     // per atom, 4 slots
    std::vector<double> s, v;           // s: nitems*3, v: natoms*3
    double box[3] = {0, 0, 0};

    std::vector<Table> bt, at, tt, nt, ot;
    std::vector<int> inbondt;           // ntype*ntype -> nb table or -1
    std::vector<int> bond_lab, angle_lab, tors_lab, oop_lab;

    // Bonded terms in the reference's evaluation order (owner ascending).
    std::vector<std::array<int, 4>> bonds, angles, tors, oops;  // member indices, -1 unused
    std::vector<int> bond_tab, angle_tab, tors_tab, oop_tab;

    // CONNECTED_TO as CSR over items.
    std::vector<int> excl_off, excl;

    // virtual sites
    std::vector<int> vs_type, vs_center, vs_off, vs_atoms;
    std::vector<double> vs_invmass;     // per VS
    std::vector<double> vs_coeff;       // per member entry: m_j / M_vs
    std::vector<int> vs_of_atom;        // per atom: owning VS or -1

    // nonbonded groups: 0 = ATOM list, 1 = BEAD list (beads then VS)
    std::vector<int> members[2];
    double rcut[2] = {0, 0}, rlist[2] = {0, 0};
    int ngroup = 1;

    int ensemble = 0, nupdate = 1, halt_drift = 1, nmts = 0, mts_quirk = 0;
    double pressure = 0, pfac = 0;      // NPT barostat (SCALEBP.f90), reduced
    double dt = 0, temp_in = 0, fac = 0;
    int numatoms = 0, numbeads = 0;
    double mktemp = 0, mktemp_atom = 0, mktemp_bead = 0;
    double temp_atom0 = 0, temp_bead0 = 0;
    // list capacity per member as a multiple of the expected neighbours within rlist
    // (a batch takes the largest of its systems')
    double list_margin = 1.5;
    // Room kept for this system's place in a batch, so that a later system of up to these sizes can take
    // the place over (ibx_assign); zero fields: the system's own sizes.  A place with a reservation
    // (reserve.items > 0) also keeps at least one padding item.
    SlotSizes reserve;
};

// What a system needs of a place in a batch (its own sizes).
SlotSizes slot_needs(const HostSystem &h);
// Every field of need within have.
bool slot_fits(const SlotSizes &have, const SlotSizes &need);
// List capacity per member (group 0) the system's density asks for (a batch takes the largest of its systems').
int list_capacity_needed(const HostSystem &h);

// Set-up of main.F90:107-234 from a descriptor (no device work).
int build_host_system(const ibx_system_desc *d, HostSystem &hs, std::string &err);

// Reference input directory -> descriptor storage (ibx_load_dir).
struct DescStorage {
    ibx_system_desc d{};
    std::vector<double> mass, s, v, tx[5], tv[5];
    std::vector<int> label, itype, nbonds, jbond, ttypes[5], toff[5], tnrow[5];
    std::vector<int> vs_type, vs_center, vs_off, vs_atoms, virtbonds, virtangles;
    int ntrj = 100, nstep = 0;
};
int read_input_dir(const char *dir, DescStorage &st, std::string &err);

// Context of already built host systems (moved in); errors as ibx_create.
int create_context(std::vector<HostSystem> &&hs, ibx_ctx **out);
// Device bytes a one-system context of h allocates at creation, plus an RDF of
// rdf_nbin bins (admission by footprint).
size_t estimate_device_bytes(const HostSystem &h, int rdf_nbin);
size_t estimate_device_bytes(const HostSystem &h, const SlotSizes &res, int rdf_nbin);
// Free device memory now (hipMemGetInfo).
int device_free_bytes(size_t *free_bytes);

// Place k of context c, whose system has been retired, taken over by h (moved in) from context step step0 on
// (a multiple of the list interval and of the RESPA factor): the place's share of every device array is
// rewritten, its record, observables, window sums and RDF start afresh, and its first lists are built; the
// other systems' data are untouched, so they continue bit for bit.  Systems without virtual sites and with
// one nonbonded group only, into a batch of such systems.
int assign_system(ibx_ctx *c, int k, HostSystem &&h, int step0);
// What place k of c holds.
SlotSizes slot_capacity(const ibx_ctx *c, int k);
// System k of c as set up on the host (null for an invalid index).
const HostSystem *host_system(const ibx_ctx *c, int k);
// Steps of one graph block of c (list interval, RESPA factor and MTS cycle); a place taken over at a multiple
// of it starts its own steps in phase with the batch's.
int batch_period(const ibx_ctx *c);

}  // namespace ibx
