// C ABI of the ibx library: context creation, device layout, per-step
// orchestration (eager and HIP-graph), observables and test hooks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ibx.h"
#include "ibx_internal.h"
#include "ibx_kernels.h"
#include "kernels_api.h"

using namespace ibx;

namespace {

thread_local std::string g_create_err;

// Graph capture on one context's stream, and device allocations, frees and
// stream/event/graph destruction of every context, are kept apart: several
// contexts may be driven from different host threads (the serving loop).
std::mutex g_capture_mu;

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) {
            std::lock_guard<std::mutex> lk(g_capture_mu);
            (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        n = (size_t)upload_count((long)count);   // never empty: kernels may read element 0 (ibx_index.h)
        std::lock_guard<std::mutex> lk(g_capture_mu);
        return hipMalloc((void **)&p, n * sizeof(T));
    }
    size_t bytes() const { return p ? n * sizeof(T) : 0; }
};

int pad_to(int x, int m) { return (x + m - 1) / m * m; }

struct GroupDev {
    int nslot_pad = 0, nwg = 0, ncell_total = 0, cap = 0, max_rows = 0;
    bool multi = false, present = false;
    DBuf<int2> nneb;
    DBuf<int> mem, mem_cell, mem_rank, cell_count, cell_start, tile_sum, tmp, perm, slot_cell, inv_perm, list, wg_sys,
        cell_sys;
    DBuf<double4> r64;
    DBuf<int4> pos4, pos4_ref;
    DBuf<double> part;   // two buffers of [kNbPart fields][workgroups] nonbonded partials
    double *part_buf(int par) { return part.p + (size_t)par * (part.n / 2); }
};

struct Timing {
    double ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

}  // namespace

// Where one place of a batch lives in the device arrays, and what it holds
// (a system assigned to the place later must fit cap).
struct SysLayout {
    SlotSizes cap;
    int item_off = 0, term_off = 0, contrib_off = 0, excl_off = 0, vsm_off = 0;
    int nb_row_off = 0, nb_tab_off = 0, nb_map_off = 0, bd_tab_off = 0, bd_row_off = 0;
    int slot_off[2] = {0, 0}, cell_off[2] = {0, 0};
};

struct ibx_ctx {
    int nsys = 0;
    std::vector<SysLayout> lay;
    int device = 0;                          // HIP device the context was created on
    std::vector<HostSystem> hs;
    std::vector<DevSys> dsys;
    std::vector<int> item_off, slot_off[2];
    int n_items = 0, item_nwg = 0, nupdate = 1;
    hipStream_t st = nullptr;
    int nb_par = 0, bd_par = 0;              // nonbonded / bonded partial buffer of the latest evaluation
    FinLaunch fin{};                         // field sums left for the next force launch
    bool fin_pending = false;
    GroupDev gr[2];
    DBuf<DevSys> sys;
    DBuf<int> wg_items, itype, label, excl_off, excl, vs_of, vs_center, vs_moff, vs_members;
    DBuf<int> term_meta, inc_off, wg_terms, nb_map, flags, step;
    DBuf<int> dmax;                  // [group][system] largest displacement^2 since the rebuild (float bits)
    DBuf<int4> term_slot;           // contribution slot of each member of a term
    DBuf<int4> term_mem;
    DBuf<Contrib> contrib;          // bonded contribution slots
    int nterms = 0, term_nwg = 0, ncontrib = 0;
    DBuf<double> s, v, mass, invmass, vs_coeff, vs_invmass, part_bd, part_mv, fsum, obs, avg, sysstate;
    DBuf<double> vs_com;             // per virtual site (SoA like s): the mass-weighted centre of its atoms
    DBuf<double> npt;                // [system][kNptState] barostat state (pressure ring, scale factor)
    bool npt_on = false;             // some system is NPT: the box rescaling runs after every move
    double *part_bd_buf(int par) { return part_bd.p + (size_t)par * (part_bd.n / 2); }   // two buffers
    DBuf<float> fnb, ftot;
    DBuf<float4> nb_rows, bd_rows;
    DBuf<double> bd_x;
    DBuf<NbTab> nb_tabs;
    DBuf<float> nb_xoff;        // x_k/bin - k per nonbonded row (non-uniform tables)
    DBuf<BdTab> bd_tabs;
    std::vector<NbTab> h_nb_tabs;
    // rdf
    DBuf<unsigned long long> rdf;
    DBuf<int> rdf_samples;           // histograms accumulated per system (counted on the device)
    DBuf<double> rdf_ivol;           // per system: the sum over the samples of 1 / volume
    int rdf_nbin = 0, rdf_every = 0;
    bool skin_tune = true;           // list-skin autotuner at every scheduled rebuild
    // force launch tiles (ibx_set_tiles): requested (-1: the tile autotuner's choice) and in use, the work
    // queues of the persistent launch, the CUs of the device, and the tuner's timings per candidate (ms)
    int tiles_req[IBX_NTILES] = {-1, -1, -1};
    ForceTiles ftiles;
    bool tiles_pending = true;
    DBuf<int> fqueue;
    int ncu = 0;
    float tile_ms[kTileCandidates] = {-1.f, -1.f, -1.f, -1.f};
    int respa = 1;                   // nonbonded force every respa steps, with weight respa
    bool mts = false;                // MTS cycle of a hybrid system (NMTS > 0)
    DBuf<float> mts_f;               // [3 slots][xyz][items] bead-group forces of the explicit steps
    DBuf<double> mts_part;           // [3 slots][9 fields][workgroups] bead-group energy/virial partials
    double rdf_dr = 0;
    // graph blocks, keyed by length, phase in the block period and the partial buffers they start from
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        int steps = 0, phase = 0;
        int par = 0, bd_par = 0;   // nb_par, bd_par when the graph was captured (and after each replay)
        uint64_t used = 0;         // last use, for eviction
    };
    static constexpr int kGraphSlots = 8;
    GraphSlot graph[kGraphSlots];
    uint64_t graph_clock = 0;
    // counters: [0] eager steps, [1] steps replayed from graphs, [2] graph launches, [3] captures
    int64_t counters[4] = {0, 0, 0, 0};
    // timing
    bool timing = false;
    Timing tm;
    std::vector<hipEvent_t> evpool;
    std::string err;
    ~ibx_ctx() {
        std::lock_guard<std::mutex> lk(g_capture_mu);
        for (auto &g : graph)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        for (auto e : evpool) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

int fail(ibx_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g_create_err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                                  \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return fail(ctx, e_ == hipErrorOutOfMemory ? IBX_E_NOMEM : IBX_E_DEVICE, "%s: %s", #expr,      \
                        hipGetErrorString(e_));                                                            \
    } while (0)

// Steps per captured graph: a whole number of list intervals and multiple-time-step cycles.
int block_period(const ibx_ctx *c) {
    auto lcm = [](int x, int y) {
        int a = x, b = y;
        while (b) { const int r = a % b; a = b; b = r; }
        return x / a * y;
    };
    int p = lcm(c->nupdate, c->respa);
    if (c->mts) p = lcm(p, 3 + c->hs[0].nmts);
    // the nonbonded partials alternate between two buffers at each evaluation,
    // so a block holds an even number of evaluations
    if ((p / c->respa) % 2) p *= 2;
    return p;
}

void drop_graph(ibx_ctx *c) {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    for (auto &g : c->graph) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = ibx_ctx::GraphSlot();
    }
}

// Copies and fills go through the context's own (non-blocking) stream, never
// the legacy null stream: another host thread may be capturing a graph on its
// context's stream meanwhile (the serving loop), and a null-stream operation
// would have to order itself against that capture.
hipError_t memcpy_st(ibx_ctx *c, void *dst, const void *src, size_t n, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, n, kind, c->st);
    return e != hipSuccess ? e : hipStreamSynchronize(c->st);
}

hipError_t memset_st(ibx_ctx *c, void *p, int v, size_t n) {
    hipError_t e = hipMemsetAsync(p, v, n, c->st);
    return e != hipSuccess ? e : hipStreamSynchronize(c->st);
}

// An empty h still allocates one element, zeroed, so kernels may read it.
template <typename T>
int upload(ibx_ctx *c, DBuf<T> &b, const std::vector<T> &h) {
    HIPCHK(c, b.alloc(h.size()));
    if (!h.empty()) HIPCHK(c, memcpy_st(c, b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    else HIPCHK(c, memset_st(c, b.p, 0, b.bytes()));
    return 0;
}

// float4 row k = {F_k, V_k, F_k+1, V_k+1}, k = 0..ndat (row ndat+1 of the table is zero)
void append_rows(std::vector<float4> &rows, const Table &t) {
    for (int k = 0; k <= t.ndat; ++k)
        rows.push_back(make_float4((float)t.F[k], (float)t.V[k], (float)t.F[k + 1], (float)t.V[k + 1]));
}

// Header of a nonbonded table in a system of working unit `unit`, and the
// offsets x_k/bin - k of its rows.  NONBONDED_FORCE.F90:66-68 takes
// NI = INT(r/bin) and alpha = (r - x_NI)/bin with the file's own abscissas;
// where those are not exactly k*bin (abscissas rounded to single precision in
// the file, for instance) the kernel subtracts the offset from alpha.
// Nonbonded rows {F_k, V_k, F_k+1 - F_k, V_k+1 - V_k}: the differences are
// formed in single precision, as a lerp between the two stored rows would.
void append_nb_rows(std::vector<float4> &rows, const Table &t) {
    for (int k = 0; k <= t.ndat; ++k) {
        const float f0 = (float)t.F[k], v0 = (float)t.V[k], f1 = (float)t.F[k + 1], v1 = (float)t.V[k + 1];
        rows.push_back(make_float4(f0, v0, f1 - f0, v1 - v0));
    }
}

NbTab nb_header(const Table &t, double unit, int row_off, std::vector<float> &xoff) {
    NbTab nt;
    nt.row_off = row_off;
    nt.ndat = t.ndat;
    nt.c = unit / t.bin;
    nt.bin = (float)t.bin;
    nt.uniform = 1;
    for (int k = 0; k <= t.ndat; ++k) {
        const double o = t.x[k] / t.bin - k;
        xoff.push_back((float)o);
        if (std::fabs(o) > 1e-9) nt.uniform = 0;
    }
    return nt;
}

// The mass-weighted centre of virtual site v's atoms (unwrapped): where DISTRIBUTE_VSFORCE.f90:27-37 puts a
// force on the site (each atom takes m_j/M of it), so where the bonded virial of NEW_FORCE.F90:73-80 counts it.
void vs_centre_of_atoms(const HostSystem &h, int v, double com[3]) {
    com[0] = com[1] = com[2] = 0.0;
    for (int q = h.vs_off[v]; q < h.vs_off[v + 1]; ++q) {
        const int a = h.vs_atoms[q];
        for (int k = 0; k < 3; ++k) com[k] += h.mass[h.itype[a]] * h.s[3 * (size_t)a + k];
    }
    for (int k = 0; k < 3; ++k) com[k] *= h.vs_invmass[v];
}

// Constants of system h in its record (the caller sets the offsets into the batch's arrays).
void set_sys_constants(const HostSystem &h, DevSys &D) {
    for (int k = 0; k < 3; ++k) {
        D.box[k] = h.box[k];
        D.boxinv[k] = 1.0 / h.box[k];
        D.box2[k] = h.box[k] / 2.0;
        D.inv_box_frac[k] = 1.0 / h.box[k];
        D.rho[k] = (float)(h.box[k] / h.box[0]);
    }
    D.unit = h.box[0] / 4294967296.0;   // fixed-point unit of the system
    D.natoms = h.natoms;
    D.nitems = h.nitems;
    D.nvs = h.nvs;
    D.hybrid = h.hybrid;
    D.ngroup = h.ngroup;
    D.dt = h.dt; D.temp_in = h.temp_in; D.fac = h.fac; D.ensemble = h.ensemble;
    D.pressure = h.pressure; D.pfac = h.pfac;
    D.mktemp = h.mktemp; D.mktemp_atom = h.mktemp_atom; D.mktemp_bead = h.mktemp_bead;
    D.volume = h.box[0] * h.box[1] * h.box[2];
    D.halt_drift = h.halt_drift < 0 ? -h.halt_drift : h.halt_drift;
    D.nmts = h.nmts; D.mts_quirk = h.mts_quirk; D.nupdate = h.nupdate;
    D.live = 1;
    D.step0 = 0;
    D.last_step = INT_MAX;
}

// Cell grid, radii and FP32 bands of group g of h in a system of fixed-point unit `unit` (the caller sets the
// offsets and member count).
void set_group_constants(const HostSystem &h, int g, double unit, DevGroup &dg) {
    const double rc = h.rcut[g], rl = h.rlist[g];
    int nc[3];
    for (int k = 0; k < 3; ++k) {
        nc[k] = (int)(cell_div(k) * h.box[k] / rl);   // cells of edge >= rlist/cell_div(k) (the reference: rlist)
        if (nc[k] < 1) nc[k] = 1;
        dg.nc[k] = nc[k];
        dg.ncmax[k] = nc[k];
    }
    dg.all_pairs = (nc[0] < min_cells(0) || nc[1] < min_cells(1) || nc[2] < min_cells(2)) ? 1 : 0;
    dg.ncell = nc[0] * nc[1] * nc[2];
    dg.rl2 = rl * rl;
    const double rl2u = (rl / unit) * (rl / unit);
    const double lband = 4e-6 + 16.0 * unit / rl;
    dg.rl2u_lo = (float)(rl2u * (1.0 - lband));
    dg.rl2u_hi = (float)(rl2u * (1.0 + lband));
    dg.rc2u = (float)((rc / unit) * (rc / unit));
    // FP32 r^2 from the fixed-point displacement is within ~2e-7 of the FP64 value, plus the
    // fixed-point rounding of the two positions (a few units over a distance rc)
    const double band = 1e-6 + 16.0 * unit / rc;
    dg.rc2 = rc * rc;
    dg.rc2u_lo = (float)((rc / unit) * (rc / unit) * (1.0 - band));
    dg.rc2u_hi = (float)((rc / unit) * (rc / unit) * (1.0 + band));
    // front run: pairs closer than rc + delta at the rebuild; the back run is scanned only once a
    // member has moved 0.49 delta (two members closing delta apart), so no pair is ever missed
    const double delta = kInnerFrac * std::max(0.0, rl - rc);
    dg.rin2u = (float)(((rc + delta) / unit) * ((rc + delta) / unit));
    dg.dmax2u = (float)((0.49 * delta / unit) * (0.49 * delta / unit));
}

int build_device(ibx_ctx *c) {
    const int S = c->nsys;
    c->item_off.resize(S + 1);
    c->slot_off[0].assign(S + 1, 0);
    c->slot_off[1].assign(S + 1, 0);
    c->dsys.assign(S, DevSys{});
    int off = 0;
    c->lay.assign(S, SysLayout());
    for (int s = 0; s < S; ++s) {
        const HostSystem &h = c->hs[s];
        c->item_off[s] = off;
        off += pad_to(std::max(h.nitems, h.reserve.items) + (h.reserve.items > 0 ? 1 : 0), kBlock);
    }
    c->item_off[S] = off;
    c->n_items = off;
    c->item_nwg = off / kBlock;
    const int N = off;

    // per-item arrays
    std::vector<double> hs_s(3 * (size_t)N, 0.0), hv(3 * (size_t)N, 0.0), hm(N, 1.0), him(N, 1.0), hvc(N, 0.0),
        hvim(N, 0.0);
    std::vector<int> hit(N, -1), hlab(N, 0), hvs(N, -1), hvcen(N, -1), hwg(c->item_nwg, 0);
    std::vector<int> hex_off(N + 1, 0), hex;
    std::vector<int> hvm_off(N + 1, 0), hvm;
    std::vector<double> hcom;         // virtual sites' centres of their atoms, when there are sites
    std::vector<int4> htm;            // term members
    std::vector<int> hmeta, hslot, hwgt;
    std::vector<BdTab> hbt;
    int nslot = 0;
    std::vector<float4> hbrows, hnrows;
    std::vector<double> hbx;
    std::vector<NbTab> hntab;
    std::vector<float> hnxoff;
    std::vector<int> hnmap;
    // contribution slots of the bonded terms each item takes part in, in term order
    std::vector<std::vector<int>> inc(N);

    for (int s = 0; s < S; ++s) {
        HostSystem &h = c->hs[s];
        DevSys &D = c->dsys[s];
        const int o = c->item_off[s];
        for (int w = o / kBlock; w < c->item_off[s + 1] / kBlock; ++w) hwg[w] = s;
        for (int i = 0; i < h.nitems; ++i) {
            for (int k = 0; k < 3; ++k) hs_s[(size_t)k * N + o + i] = h.s[3 * (size_t)i + k];
            if (i < h.natoms)
                for (int k = 0; k < 3; ++k) hv[(size_t)k * N + o + i] = h.v[3 * (size_t)i + k];
            hit[o + i] = h.itype[i];
            hlab[o + i] = h.type_label[i];
            hm[o + i] = h.mass[h.itype[i]];
            him[o + i] = h.invmass[h.itype[i]];
        }
        // exclusions in global indices
        for (int i = 0; i < h.nitems; ++i)
            for (int q = h.excl_off[i]; q < h.excl_off[i + 1]; ++q) hex.push_back(o + h.excl[q]);
        // virtual sites
        if (h.nvs > 0 && hcom.empty()) hcom.assign(3 * (size_t)N, 0.0);
        for (int v = 0; v < h.nvs; ++v) {
            const int vi = o + h.natoms + v;
            hvcen[vi] = h.vs_center[v] >= 0 ? o + h.vs_center[v] : -1;
            hvim[vi] = h.vs_invmass[v];
            double com[3];
            vs_centre_of_atoms(h, v, com);
            for (int k = 0; k < 3; ++k) hcom[(size_t)k * N + vi] = com[k];
            for (int q = h.vs_off[v]; q < h.vs_off[v + 1]; ++q) {
                hvs[o + h.vs_atoms[q]] = vi;
                hvc[o + h.vs_atoms[q]] = h.vs_coeff[q];
            }
        }
        // bonded tables of this system
        auto add_tabs = [&](const std::vector<Table> &tabs, const std::vector<int> &labs) {
            int base = (int)hbt.size();
            for (size_t t = 0; t < tabs.size(); ++t) {
                BdTab b;
                b.row_off = (int)hbrows.size();
                b.ndat = tabs[t].ndat;
                b.x0 = tabs[t].x0;
                b.bin = tabs[t].bin;
                b.inv_bin = 1.0 / tabs[t].bin;
                b.pad = 0;
                b.label = labs.empty() ? 1 : labs[t];
                hbt.push_back(b);
                append_rows(hbrows, tabs[t]);
                for (int k = 0; k <= tabs[t].ndat; ++k) hbx.push_back(tabs[t].x[k]);
            }
            return base;
        };
        SysLayout &L = c->lay[s];
        L.bd_tab_off = (int)hbt.size();
        L.bd_row_off = (int)hbrows.size();
        int bbase = add_tabs(h.bt, h.bond_lab), abase = add_tabs(h.at, h.angle_lab);
        int tbase = add_tabs(h.tt, h.tors_lab), obase = add_tabs(h.ot, h.oop_lab);
        // reserved room for a later system's tables (zero headers and rows, never read)
        L.cap.bd_tabs = std::max((int)hbt.size() - L.bd_tab_off, h.reserve.bd_tabs);
        L.cap.bd_rows = std::max((int)hbrows.size() - L.bd_row_off, h.reserve.bd_rows);
        hbt.resize(L.bd_tab_off + L.cap.bd_tabs, BdTab{});
        hbrows.resize(L.bd_row_off + L.cap.bd_rows, make_float4(0.f, 0.f, 0.f, 0.f));
        hbx.resize(L.bd_row_off + L.cap.bd_rows, 0.0);
        // terms of this system, padded to whole workgroups (one system per workgroup)
        const int term0 = (int)hmeta.size();
        auto add_term = [&](int kind, int tab, const std::array<int, 4> &m, int nm) {
            int4 g4 = make_int4(o + m[0], o + m[1], nm > 2 ? o + m[2] : -1, nm > 3 ? o + m[3] : -1);
            htm.push_back(g4);
            int vsmask = 0;   // members that are virtual sites (terms of the virtangles file)
            for (int r = 0; r < nm; ++r)
                if (m[r] >= h.natoms) vsmask |= 1 << r;
            hmeta.push_back(kind | (tab << 2) | (vsmask << kTermVsShift));
            hslot.push_back(nslot);
            for (int r = 0; r < nm; ++r) inc[o + m[r]].push_back(nslot + r);
            nslot += nm;
        };
        for (size_t q = 0; q < h.bonds.size(); ++q) add_term(KBOND, bbase + h.bond_tab[q], h.bonds[q], 2);
        for (size_t q = 0; q < h.angles.size(); ++q) add_term(KANGLE, abase + h.angle_tab[q], h.angles[q], 3);
        for (size_t q = 0; q < h.tors.size(); ++q) add_term(KTORS, tbase + h.tors_tab[q], h.tors[q], 4);
        for (size_t q = 0; q < h.oops.size(); ++q) add_term(KOOP, obase + h.oop_tab[q], h.oops[q], 4);
        int nt_pad = pad_to(std::max(1, std::max((int)hmeta.size() - term0, h.reserve.terms)), kBlock);
        while ((int)hmeta.size() < term0 + nt_pad) {
            htm.push_back(make_int4(-1, -1, -1, -1));
            hmeta.push_back(-1);
            hslot.push_back(0);
        }
        D.term_wg_off = term0 / kBlock;
        D.term_nwg = nt_pad / kBlock;
        L.term_off = term0;
        L.cap.terms = nt_pad;
        for (int w = 0; w < D.term_nwg; ++w) hwgt.push_back(s);
        // nonbonded tables: fixed-point unit of this system
        const double unit = h.box[0] / 4294967296.0;
        D.nb_row_off = (int)hnrows.size();
        D.nb_tab_off = (int)hntab.size();
        D.nb_ntab = (int)h.nt.size();
        if (D.nb_ntab > kMaxNbTabs) return fail(c, IBX_E_INVALID, "more than %d nonbonded tables", kMaxNbTabs);
        for (auto &t : h.nt) {
            hntab.push_back(nb_header(t, unit, (int)hnrows.size(), hnxoff));
            append_nb_rows(hnrows, t);
        }
        D.nb_nrow = (int)hnrows.size() - D.nb_row_off;
        D.ntype = h.ntype;
        D.nbmap_off = (int)hnmap.size();
        if (h.nt.size() > 1 && h.ntype > kMaxNbTypes)
            return fail(c, IBX_E_INVALID, "more than %d types with several nonbonded tables", kMaxNbTypes);
        for (int q = 0; q < h.ntype * h.ntype; ++q) hnmap.push_back(h.inbondt[q] < 0 ? -1 : D.nb_tab_off + h.inbondt[q]);
        L.nb_row_off = D.nb_row_off;
        L.nb_tab_off = D.nb_tab_off;
        L.nb_map_off = D.nbmap_off;
        L.cap.nb_rows = std::max(D.nb_nrow, h.reserve.nb_rows);
        L.cap.nb_tabs = std::max(D.nb_ntab, h.reserve.nb_tabs);
        L.cap.nb_map = std::max(h.ntype * h.ntype, h.reserve.nb_map);
        hnrows.resize(L.nb_row_off + L.cap.nb_rows, make_float4(0.f, 0.f, 0.f, 0.f));
        hnxoff.resize(L.nb_row_off + L.cap.nb_rows, 0.f);
        hntab.resize(L.nb_tab_off + L.cap.nb_tabs, NbTab{});
        hnmap.resize(L.nb_map_off + L.cap.nb_map, -1);
        set_sys_constants(h, D);
        D.item_off = o;
        D.nitems_pad = c->item_off[s + 1] - o;
        D.item_wg_off = o / kBlock;
        D.item_nwg = D.nitems_pad / kBlock;
    }
    for (int i = 0; i < N; ++i) {
        hex_off[i + 1] = (int)hex.size();   // placeholder, rebuilt below
    }
    // exclusion CSR in global item order
    {
        hex.clear();
        for (int s = 0; s < S; ++s) {
            HostSystem &h = c->hs[s];
            const int o = c->item_off[s];
            SysLayout &L = c->lay[s];
            L.excl_off = (int)hex.size();
            for (int i = o; i < c->item_off[s + 1]; ++i) {
                hex_off[i] = (int)hex.size();
                int li = i - o;
                if (li < h.nitems)
                    for (int q = h.excl_off[li]; q < h.excl_off[li + 1]; ++q) hex.push_back(o + h.excl[q]);
            }
            // reserved room: entries of the last (padding) item, never read
            while ((int)hex.size() < L.excl_off + h.reserve.excl) hex.push_back(o);
            L.cap.excl = (int)hex.size() - L.excl_off;
        }
        hex_off[N] = (int)hex.size();
    }
    // VS member CSR per VS item
    for (int i = 0; i < N; ++i) hvm_off[i + 1] = hvm_off[i];
    {
        hvm.clear();
        std::vector<int> cnt(N + 1, 0);
        for (int s = 0; s < S; ++s) {
            HostSystem &h = c->hs[s];
            const int o = c->item_off[s];
            c->lay[s].vsm_off = (int)hvm.size();
            for (int i = o; i < c->item_off[s + 1]; ++i) {
                hvm_off[i] = (int)hvm.size();
                int li = i - o;
                if (li >= h.natoms && li < h.nitems) {
                    int v = li - h.natoms;
                    for (int q = h.vs_off[v]; q < h.vs_off[v + 1]; ++q) hvm.push_back(o + h.vs_atoms[q]);
                }
            }
        }
        hvm_off[N] = (int)hvm.size();
    }
    // contribution slots renumbered so each item's slots are one contiguous range, in term order
    std::vector<int> hinc_off(N + 1, 0), renum((size_t)std::max(1, nslot), 0);
    for (int s = 0; s < S; ++s) {
        const int o = c->item_off[s], e = c->item_off[s + 1];
        c->lay[s].contrib_off = hinc_off[o];
        for (int i = o; i < e; ++i) {
            hinc_off[i + 1] = hinc_off[i] + (int)inc[i].size();
            for (size_t q = 0; q < inc[i].size(); ++q) renum[inc[i][q]] = hinc_off[i] + (int)q;
        }
        // reserved room: slots of the last (padding) item, never gathered
        hinc_off[e] = std::max(hinc_off[e], c->lay[s].contrib_off + c->hs[s].reserve.contrib);
        c->lay[s].cap.contrib = hinc_off[e] - c->lay[s].contrib_off;
    }
    std::vector<int4> hslot4(hslot.size());
    for (size_t t = 0; t < hslot.size(); ++t) {
        const int4 m = htm[t];
        const int b = hslot[t];
        hslot4[t] = make_int4(m.x >= 0 ? renum[b] : 0, m.y >= 0 ? renum[b + 1] : 0, m.z >= 0 ? renum[b + 2] : 0,
                              m.w >= 0 ? renum[b + 3] : 0);
    }
    c->nterms = (int)hmeta.size();
    c->term_nwg = c->nterms / kBlock;
    c->ncontrib = std::max(1, hinc_off[N]);
    for (int s = 0; s < S; ++s) {
        c->lay[s].item_off = c->item_off[s];
        c->lay[s].cap.items = c->item_off[s + 1] - c->item_off[s];
    }

    int rc;
    if ((rc = upload(c, c->s, hs_s)) || (rc = upload(c, c->v, hv)) || (rc = upload(c, c->mass, hm)) ||
        (rc = upload(c, c->invmass, him)) || (rc = upload(c, c->itype, hit)) || (rc = upload(c, c->label, hlab)) ||
        (rc = upload(c, c->wg_items, hwg)) || (rc = upload(c, c->excl_off, hex_off)) || (rc = upload(c, c->excl, hex)) ||
        (rc = upload(c, c->vs_of, hvs)) || (rc = upload(c, c->vs_coeff, hvc)) || (rc = upload(c, c->vs_center, hvcen)) ||
        (rc = upload(c, c->vs_invmass, hvim)) || (rc = upload(c, c->vs_moff, hvm_off)) ||
        (rc = upload(c, c->vs_members, hvm)) || (rc = upload(c, c->inc_off, hinc_off)) ||
        (rc = upload(c, c->term_mem, htm)) || (rc = upload(c, c->term_meta, hmeta)) ||
        (rc = upload(c, c->term_slot, hslot4)) || (rc = upload(c, c->wg_terms, hwgt)) ||
        (rc = upload(c, c->bd_tabs, hbt)) || (rc = upload(c, c->bd_rows, hbrows)) || (rc = upload(c, c->bd_x, hbx)) ||
        (rc = upload(c, c->nb_rows, hnrows)) || (rc = upload(c, c->nb_tabs, hntab)) || (rc = upload(c, c->nb_map, hnmap)) ||
        (rc = upload(c, c->nb_xoff, hnxoff)) || (rc = upload(c, c->vs_com, hcom)))
        return rc;
    c->h_nb_tabs = hntab;
    HIPCHK(c, c->fnb.alloc(3 * (size_t)N));
    HIPCHK(c, c->contrib.alloc((size_t)c->ncontrib));
    HIPCHK(c, memset_st(c, c->contrib.p, 0, c->contrib.bytes()));
    HIPCHK(c, c->ftot.alloc(3 * (size_t)N));
    HIPCHK(c, memset_st(c, c->fnb.p, 0, c->fnb.bytes()));
    HIPCHK(c, memset_st(c, c->ftot.p, 0, c->ftot.bytes()));
    HIPCHK(c, c->part_bd.alloc(2 * (size_t)c->term_nwg * kBdPart));
    HIPCHK(c, c->part_mv.alloc((size_t)c->item_nwg * kMvPart));
    HIPCHK(c, memset_st(c, c->part_bd.p, 0, c->part_bd.bytes()));
    HIPCHK(c, memset_st(c, c->part_mv.p, 0, c->part_mv.bytes()));
    HIPCHK(c, c->fsum.alloc((size_t)S * kFields));
    HIPCHK(c, c->obs.alloc((size_t)S * kObs));
    HIPCHK(c, c->avg.alloc((size_t)S * kObs));
    HIPCHK(c, c->sysstate.alloc((size_t)S * 8));
    HIPCHK(c, memset_st(c, c->obs.p, 0, c->obs.bytes()));
    HIPCHK(c, memset_st(c, c->avg.p, 0, c->avg.bytes()));
    std::vector<double> st(8 * (size_t)S, 0.0);
    for (int s = 0; s < S; ++s) {
        st[8 * s] = c->hs[s].temp_atom0;
        st[8 * s + 1] = c->hs[s].temp_bead0;
        st[8 * s + 5] = 1.0;   // the next move advances the system
    }
    if ((rc = upload(c, c->sysstate, st))) return rc;
    HIPCHK(c, c->npt.alloc((size_t)S * kNptState));
    HIPCHK(c, memset_st(c, c->npt.p, 0, c->npt.bytes()));
    c->npt_on = false;
    for (int s = 0; s < S; ++s) c->npt_on = c->npt_on || c->hs[s].ensemble == 2;
    HIPCHK(c, c->dmax.alloc(2 * (size_t)S));
    {
        std::vector<int> inf(2 * (size_t)S, 0x7f800000);   // no bound until the first rebuild
        if ((rc = upload(c, c->dmax, inf))) return rc;
    }
    HIPCHK(c, c->flags.alloc(8));
    HIPCHK(c, memset_st(c, c->flags.p, 0, c->flags.bytes()));
    HIPCHK(c, c->fqueue.alloc(kQueueInts));
    HIPCHK(c, memset_st(c, c->fqueue.p, 0, c->fqueue.bytes()));
    c->ftiles.queue = c->fqueue.p;
    HIPCHK(c, c->step.alloc(1));
    HIPCHK(c, memset_st(c, c->step.p, 0, c->step.bytes()));
    for (DevSys &D : c->dsys) D.step_ctr = c->step.p;

    // groups
    for (int g = 0; g < 2; ++g) {
        GroupDev &G = c->gr[g];
        std::vector<int> mem, wg, cellsys;
        int slot = 0, cell = 0, maxrows = 0;
        bool any = false, multi = false;
        double max_dens_vol = 0.0, margin = 1.5;
        for (int s = 0; s < S; ++s) {
            HostSystem &h = c->hs[s];
            DevSys &D = c->dsys[s];
            DevGroup &dg = D.g[g];
            const int nm = (g < h.ngroup) ? (int)h.members[g].size() : 0;
            dg.slot_off = slot;
            dg.nmem = nm;
            dg.nslot_pad = pad_to(std::max(nm, h.reserve.members[g]), kBlock);
            dg.mem_off = slot;
            dg.wg_off = slot / kBlock;
            dg.nwg = dg.nslot_pad / kBlock;
            c->lay[s].slot_off[g] = slot;
            c->lay[s].cap.members[g] = dg.nslot_pad;
            for (int k = 0; k < dg.nslot_pad; ++k) mem.push_back(k < nm ? c->item_off[s] + h.members[g][k] : -1);
            for (int w = 0; w < dg.nwg; ++w) wg.push_back(s);
            slot += dg.nslot_pad;
            set_group_constants(h, g, D.unit, dg);
            dg.cell_off = cell;
            const int ncell_res = std::max(dg.ncell, h.reserve.cells[g]);   // reserved cells stay empty
            c->lay[s].cell_off[g] = cell;
            c->lay[s].cap.cells[g] = ncell_res;
            for (int q = 0; q < ncell_res; ++q) cellsys.push_back(s);
            cell += ncell_res;
            if (nm > 0) {
                any = true;
                const double rl = h.rlist[g];
                double vol = h.box[0] * h.box[1] * h.box[2];
                double nbr = 4.0 / 3.0 * M_PI * rl * rl * rl * nm / vol;
                max_dens_vol = std::max(max_dens_vol, nbr);
                margin = std::max(margin, h.list_margin);
                maxrows = std::max(maxrows, c->lay[s].cap.nb_rows);   // rows a later system may bring too
                multi = multi || c->lay[s].cap.nb_tabs > 1;
            }
        }
        G.present = any;
        G.nslot_pad = slot;
        G.nwg = slot / kBlock;
        G.ncell_total = cell;
        G.multi = multi;
        G.max_rows = std::min(maxrows, kLdsRows);
        if (maxrows > kLdsRows) G.max_rows = 0;
        if (!any) continue;
        G.cap = pad_to((int)(margin * max_dens_vol) + 64, 8);
        std::vector<int> perm_init(slot, -1);
        if ((rc = upload(c, G.mem, mem)) || (rc = upload(c, G.wg_sys, wg)) || (rc = upload(c, G.cell_sys, cellsys)) ||
            (rc = upload(c, G.perm, perm_init)))
            return rc;
        HIPCHK(c, G.mem_cell.alloc(slot));
        HIPCHK(c, G.mem_rank.alloc(slot));
        HIPCHK(c, G.cell_count.alloc(cell));
        HIPCHK(c, memset_st(c, G.cell_count.p, 0, G.cell_count.bytes()));   // each scan clears it for the next rebuild
        HIPCHK(c, G.cell_start.alloc(cell + 1));
        HIPCHK(c, G.tile_sum.alloc((cell + 1 + kBlock - 1) / kBlock));
        HIPCHK(c, G.tmp.alloc(slot));
        HIPCHK(c, G.slot_cell.alloc(slot));
        HIPCHK(c, G.r64.alloc(slot));
        HIPCHK(c, G.pos4.alloc(slot));
        HIPCHK(c, G.pos4_ref.alloc(slot));
        HIPCHK(c, G.nneb.alloc(slot));
        HIPCHK(c, memset_st(c, G.nneb.p, 0, G.nneb.bytes()));   // padding slots keep empty rows
        HIPCHK(c, G.part.alloc(2 * (size_t)G.nwg * (kBlock / kNbAtomsPerBlock) * kNbPart));
        HIPCHK(c, memset_st(c, G.part.p, 0, G.part.bytes()));
        HIPCHK(c, G.list.alloc((size_t)slot * G.cap));
        std::vector<int> inv(N, -1);
        if ((rc = upload(c, G.inv_perm, inv))) return rc;
    }
    if ((rc = upload(c, c->sys, c->dsys))) return rc;
    return 0;
}

// Device bytes a context holds (every buffer it allocated).
size_t device_bytes(const ibx_ctx *c) {
    size_t b = c->sys.bytes() + c->wg_items.bytes() + c->itype.bytes() + c->label.bytes() + c->excl_off.bytes() +
               c->excl.bytes() + c->vs_of.bytes() + c->vs_center.bytes() + c->vs_moff.bytes() + c->vs_members.bytes() +
               c->term_meta.bytes() + c->inc_off.bytes() + c->wg_terms.bytes() + c->nb_map.bytes() + c->flags.bytes() + c->fqueue.bytes() +
               c->step.bytes() + c->dmax.bytes() + c->term_slot.bytes() + c->term_mem.bytes() + c->contrib.bytes() +
               c->s.bytes() + c->v.bytes() + c->mass.bytes() + c->invmass.bytes() + c->vs_coeff.bytes() +
               c->vs_invmass.bytes() + c->part_bd.bytes() + c->part_mv.bytes() + c->fsum.bytes() + c->obs.bytes() +
               c->avg.bytes() + c->sysstate.bytes() + c->fnb.bytes() + c->ftot.bytes() + c->nb_rows.bytes() +
               c->bd_rows.bytes() + c->bd_x.bytes() + c->nb_tabs.bytes() + c->nb_xoff.bytes() + c->bd_tabs.bytes() +
               c->rdf.bytes() + c->rdf_samples.bytes() + c->rdf_ivol.bytes() + c->mts_f.bytes() + c->mts_part.bytes() + c->vs_com.bytes() + c->npt.bytes();
    for (const GroupDev &G : c->gr)
        b += G.nneb.bytes() + G.mem.bytes() + G.mem_cell.bytes() + G.cell_count.bytes() + G.cell_start.bytes() +
             G.mem_rank.bytes() + G.tile_sum.bytes() + G.tmp.bytes() + G.perm.bytes() + G.slot_cell.bytes() + G.inv_perm.bytes() +
             G.list.bytes() + G.wg_sys.bytes() + G.cell_sys.bytes() + G.r64.bytes() + G.pos4.bytes() +
             G.pos4_ref.bytes() + G.part.bytes();
    return b;
}

}  // namespace

// What h needs of a place in a batch.
SlotSizes ibx::slot_needs(const HostSystem &h) {
    SlotSizes n;
    n.items = h.nitems;
    n.terms = (int)(h.bonds.size() + h.angles.size() + h.tors.size() + h.oops.size());
    n.contrib = (int)(2 * h.bonds.size() + 3 * h.angles.size() + 4 * (h.tors.size() + h.oops.size()));
    n.excl = (int)h.excl.size();
    for (int g = 0; g < 2; ++g) {
        n.members[g] = g < h.ngroup ? (int)h.members[g].size() : 0;
        if (n.members[g] == 0) continue;
        int cells = 1;
        for (int k = 0; k < 3; ++k) cells *= std::max(1, (int)(cell_div(k) * h.box[k] / h.rlist[g]));
        n.cells[g] = cells;
    }
    for (const Table &t : h.nt) n.nb_rows += t.ndat + 1;
    n.nb_tabs = (int)h.nt.size();
    n.nb_map = h.ntype * h.ntype;
    for (const auto *v : {&h.bt, &h.at, &h.tt, &h.ot})
        for (const Table &t : *v) n.bd_rows += t.ndat + 1;
    n.bd_tabs = (int)(h.bt.size() + h.at.size() + h.tt.size() + h.ot.size());
    return n;
}

int ibx::list_capacity_needed(const HostSystem &h) {
    const int nm = h.members[0].empty() ? 0 : (int)h.members[0].size();
    const double rl = h.rlist[0], vol = h.box[0] * h.box[1] * h.box[2];
    const double nbr = 4.0 / 3.0 * M_PI * rl * rl * rl * nm / vol;
    return (int)(std::max(1.5, h.list_margin) * nbr) + 64;
}

bool ibx::slot_fits(const SlotSizes &have, const SlotSizes &need) {
    bool ok = need.items <= have.items && need.terms <= have.terms && need.contrib <= have.contrib &&
              need.excl <= have.excl && need.nb_rows <= have.nb_rows && need.nb_tabs <= have.nb_tabs &&
              need.nb_map <= have.nb_map && need.bd_tabs <= have.bd_tabs && need.bd_rows <= have.bd_rows;
    for (int g = 0; g < 2; ++g) ok = ok && need.members[g] <= have.members[g] && need.cells[g] <= have.cells[g];
    return ok;
}

// Device bytes a one-system context of h would allocate at creation (the sizes
// build_device computes, its reservation included), plus an RDF of rdf_nbin bins;
// the second form with reservation res in place of h's own.
size_t ibx::estimate_device_bytes(const HostSystem &h, int rdf_nbin) { return estimate_device_bytes(h, h.reserve, rdf_nbin); }

size_t ibx::estimate_device_bytes(const HostSystem &h, const SlotSizes &res, int rdf_nbin) {
    auto at_least1 = [](size_t n) { return n ? n : (size_t)1; };
    const SlotSizes need = slot_needs(h);
    auto cap = [](int a, int b) { return (size_t)std::max(a, b); };
    const size_t N = (size_t)pad_to(std::max(h.nitems, res.items) + (res.items > 0 ? 1 : 0), kBlock);
    const size_t nterm_pad = (size_t)pad_to(std::max(1, std::max(need.terms, res.terms)), kBlock);
    const size_t members = cap(need.contrib, res.contrib);
    const size_t bd_rows = cap(need.bd_rows, res.bd_rows), nb_rows = cap(need.nb_rows, res.nb_rows);
    size_t b = sizeof(DevSys) + (N / kBlock) * 4                 // sys, wg_items
               + N * (4 + 4 + 4 + 4 + 8 + 8 + 8 + 8 + 24 + 24 + 12 + 12)   // itype label vs_of vs_center mass
                                                                  // invmass vs_coeff vs_invmass s v fnb ftot
               + 3 * (N + 1) * 4                                  // excl_off vs_moff inc_off
               + at_least1(cap(need.excl, res.excl)) * 4 + at_least1(h.vs_atoms.size()) * 4
               + nterm_pad * (4 + 16 + 16) + (nterm_pad / kBlock) * 4   // meta, members, slots, wg_terms
               + at_least1(members) * sizeof(Contrib)             // contributions
               + 2 * (nterm_pad / kBlock) * kBdPart * 8 + (N / kBlock) * kMvPart * 8
               + (size_t)(kFields + 2 * kObs + 8 + kNptState) * 8 + 2 * 4 + 8 * 4 + 4   // fsum obs avg sysstate npt dmax flags
                                                                       // step
               + kQueueInts * 4                                       // work queues of the persistent force launch
               + at_least1(bd_rows) * (16 + 8) + at_least1(cap(need.bd_tabs, res.bd_tabs)) * sizeof(BdTab)
               + at_least1(nb_rows) * (16 + 4) + at_least1(cap(need.nb_tabs, res.nb_tabs)) * sizeof(NbTab) +
               at_least1(cap(need.nb_map, res.nb_map)) * 4
               + (h.nvs > 0 ? 24 * N : 8);                          // vs_com
    if (rdf_nbin > 0) b += (size_t)rdf_nbin * 8 + 4 + 8;   // histogram, samples, sum of 1/volume
    if (h.hybrid && h.nmts > 0 && h.ngroup > 1 && !h.members[1].empty())   // MTS slots
        b += 9 * N * 4 + 27 * (size_t)(pad_to((int)h.members[1].size(), kBlock) / kBlock) * (kBlock / kNbAtomsPerBlock) * 8;
    for (int g = 0; g < 2; ++g) {
        const int nm = need.members[g];
        if (nm == 0) continue;   // an absent group allocates nothing
        const size_t slots = (size_t)pad_to(std::max(nm, res.members[g]), kBlock), nwg = slots / kBlock;
        const size_t ncell = cap(need.cells[g], res.cells[g]);
        const double nbr = 4.0 / 3.0 * M_PI * h.rlist[g] * h.rlist[g] * h.rlist[g] * nm /
                           (h.box[0] * h.box[1] * h.box[2]);
        const size_t lcap = (size_t)pad_to((int)(std::max(1.5, h.list_margin) * nbr) + 64, 8);
        b += slots * (4 * 6 + 32 + 16 + 16 + 8)                  // mem mem_cell mem_rank tmp perm slot_cell r64 pos4 pos4_ref nneb
             + N * 4 + nwg * 4 + ncell * 12 + 4                   // inv_perm, wg_sys, cell arrays (one more start)
             + (ncell + 1 + kBlock - 1) / kBlock * 4              // scan tile sums
             + 2 * nwg * (kBlock / kNbAtomsPerBlock) * kNbPart * 8 + slots * lcap * 4;
    }
    return b;
}

namespace {

RebuildArgs rebuild_args(ibx_ctx *c, int g) {
    GroupDev &G = c->gr[g];
    RebuildArgs a;
    a.sys = c->sys.p; a.wg_sys = G.wg_sys.p; a.cell_sys = G.cell_sys.p;
    a.g = g; a.nwg = G.nwg; a.ncell_total = G.ncell_total; a.nitems_total = c->n_items; a.cap = G.cap;
    a.cubic = true;
    for (const DevSys &D : c->dsys) a.cubic = a.cubic && D.rho[1] == 1.0f && D.rho[2] == 1.0f;
    a.mem = G.mem.p; a.s = c->s.p; a.itype = c->itype.p; a.label = c->label.p;
    a.mem_cell = G.mem_cell.p; a.cell_count = G.cell_count.p; a.cell_start = G.cell_start.p;
    a.mem_rank = G.mem_rank.p; a.tile_sum = G.tile_sum.p; a.tmp = G.tmp.p; a.perm = G.perm.p; a.slot_cell = G.slot_cell.p;
    a.r64 = G.r64.p; a.pos4 = G.pos4.p; a.inv_perm = G.inv_perm.p; a.pos4_ref = G.pos4_ref.p;
    a.dmax = c->dmax.p + (size_t)g * c->nsys; a.nsys = c->nsys;
    a.tune = -1; a.nmove = c->nupdate; a.tune_sys = c->sys.p; a.dmax_all = c->dmax.p;
    a.excl_off = c->excl_off.p; a.excl = c->excl.p; a.list = G.list.p; a.nneb = G.nneb.p; a.flags = c->flags.p;
    a.only = -1;
    a.rdf_hist = nullptr;
    a.rdf_samples = nullptr;
    a.rdf_ivol = nullptr;
    a.rdf_nbin = 0;
    a.rdf_inv_dr = 0.f;
    return a;
}

// tune: a rebuild of the step schedule, after a full interval of moves (the
// list-skin autotuner reads the displacement that interval reached); with the
// RDF on, the group-0 list build also adds its listed pairs to each system's
// histogram (one sample a build).
void enqueue_rebuild(ibx_ctx *c, bool tune = false) {
    const bool rdf = tune && c->rdf_nbin > 0 && c->rdf_every > 0;
    for (int g = 0; g < 2; ++g) {
        if (!c->gr[g].present) continue;
        RebuildArgs a = rebuild_args(c, g);
        if (tune) a.tune = c->skin_tune ? 1 : 0;   // with the first group's cell count, for both groups
        tune = false;
        if (rdf && g == 0) {
            a.rdf_hist = c->rdf.p;
            a.rdf_samples = c->rdf_samples.p;
            a.rdf_ivol = c->rdf_ivol.p;
            a.rdf_nbin = c->rdf_nbin;
            a.rdf_inv_dr = (float)(1.0 / c->rdf_dr);
        }
        launch_rebuild(c->st, a);
    }
}

NbArgs nb_args(ibx_ctx *c, int g) {
    GroupDev &G = c->gr[g];
    NbArgs a;
    a.sys = c->sys.p; a.wg_sys = G.wg_sys.p; a.g = g; a.nwg = G.nwg; a.cap = G.cap; a.nitems_total = c->n_items;
    a.max_rows_lds = G.max_rows; a.multi = G.multi; a.perm = G.perm.p; a.pos4 = G.pos4.p; a.list = G.list.p;
    a.nneb = G.nneb.p; a.rows = c->nb_rows.p; a.xoff = c->nb_xoff.p; a.tabs = c->nb_tabs.p; a.map = c->nb_map.p;
    a.fnb = c->fnb.p; a.part = G.part_buf(c->nb_par); a.s = c->s.p; a.dmax = c->dmax.p; a.nsys = c->nsys;
    a.cubic = true;
    for (const DevSys &D : c->dsys) a.cubic = a.cubic && D.rho[1] == 1.0f && D.rho[2] == 1.0f;
    return a;
}

// The nonbonded force of groups g0..1, into the partial buffer nb_par (the
// caller has switched it).  mode 0: the force phase of step `step`; with the
// MTS cycle (hybrid, NMTS > 0, MTS_LOOP.f90:71-224) the bead group is
// evaluated and saved on explicit steps and extrapolated on the others.
// mode 1: an explicit evaluation outside the step loop.
void enqueue_nb_groups(ibx_ctx *c, int mode, int step, int g0) {
    const bool mts = mode == 0 && c->mts;
    const int nmts = c->hs[0].nmts;
    const int mod = mts ? step % (3 + nmts) : 1;
    for (int g = g0; g < 2; ++g) {
        GroupDev &G = c->gr[g];
        if (!G.present) continue;
        const int npart = G.nwg * (kBlock / kNbAtomsPerBlock);
        if (g == 1 && mts && (mod < 1 || mod > 3)) {
            // MTS_APPROX: I = MOD(STEP, 3+NMTS), which is 0 on the last step of a cycle in the reference
            const double I = (mod == 0 && !c->hs[0].mts_quirk) ? (double)(3 + nmts) : (double)mod;
            launch_mts(c->st, false, c->fnb.p, c->label.p, c->n_items, G.part_buf(c->nb_par), npart, c->mts_f.p,
                       c->mts_part.p, 0, I);
            continue;
        }
        launch_nonbonded(c->st, nb_args(c, g));
        if (g == 1 && mts)   // MTS_SAVEFORCE, MTS_SAVEAUXS into slot MOD - 1
            launch_mts(c->st, true, c->fnb.p, c->label.p, c->n_items, G.part_buf(c->nb_par), npart, c->mts_f.p,
                       c->mts_part.p, mod - 1, 0.0);
    }
}

// Each evaluation writes the other of the two nonbonded partial buffers (and
// each bonded evaluation the other of the two bonded ones), so the field sums
// of the previous step can read theirs while the next step's forces run.
void enqueue_forces(ibx_ctx *c, int mode, int step = 0) {
    c->nb_par ^= 1;
    enqueue_nb_groups(c, mode, step, 0);
}

TermLaunch term_launch(const ibx_ctx *c) {
    return TermLaunch{c->sys.p, c->term_mem.p, c->term_meta.p, c->term_slot.p, c->bd_tabs.p, c->bd_rows.p, c->bd_x.p,
                      c->vs_com.p};
}

void enqueue_terms(ibx_ctx *c) {
    c->bd_par ^= 1;
    launch_terms(c->st, c->term_nwg, c->wg_terms.p, term_launch(c), c->s.p, c->n_items, c->nterms, c->contrib.p,
                 c->ncontrib, c->part_bd_buf(c->bd_par), c->flags.p);
}

// The field sums of the latest evaluation (nonbonded, bonded and integrator partials).
FinLaunch fin_launch(ibx_ctx *c, int mode) {
    constexpr int per = kBlock / kNbAtomsPerBlock;
    GroupDev &G1 = c->gr[1].present ? c->gr[1] : c->gr[0];
    FinLaunch f;
    f.sys = c->sys.p; f.part_nb0 = c->gr[0].part_buf(c->nb_par); f.part_nb1 = G1.part_buf(c->nb_par);
    f.part_bd = c->part_bd_buf(c->bd_par); f.part_mv = c->part_mv.p;
    f.snb0 = per * c->gr[0].nwg; f.snb1 = per * G1.nwg; f.sbd = c->term_nwg; f.smv = c->item_nwg;
    f.fsum = c->fsum.p; f.obs = c->obs.p; f.avg = c->avg.p; f.sysstate = c->sysstate.p;
    f.step = c->step.p; f.mode = mode; f.nsys = c->nsys;
    f.npt = c->npt.p;
    return f;
}

// The field sums a step left for the next step's force launch, enqueued on their own.
void flush_fin(ibx_ctx *c) {
    if (!c->fin_pending) return;
    launch_finalize(c->st, c->fin);
    c->fin_pending = false;
}

// Integration and reductions after the forces (the bonded terms must be enqueued before).
// defer: the field sums are left to the next step's force launch (flush_fin
// enqueues them when no step follows).
void enqueue_rest(ibx_ctx *c, int mode, std::vector<hipEvent_t> *ev, float nb_weight = 1.0f, bool defer = false) {
    MoveLaunch m;
    m.nb_weight = mode == 1 ? 1.0f : nb_weight;
    m.s = c->s.p; m.v = c->v.p; m.fnb = c->fnb.p; m.contrib = c->contrib.p; m.inc_off = c->inc_off.p;
    m.ncontrib = c->ncontrib; m.ftot = c->ftot.p; m.itype = c->itype.p;
    m.label = c->label.p; m.mass = c->mass.p; m.invmass = c->invmass.p; m.vs_of = c->vs_of.p;
    m.vs_coeff = c->vs_coeff.p;
    for (int g = 0; g < 2; ++g) {
        m.pos4[g] = c->gr[g].present ? c->gr[g].pos4.p : nullptr;
        m.pos4_ref[g] = c->gr[g].present ? c->gr[g].pos4_ref.p : nullptr;
        m.inv_perm[g] = c->gr[g].present ? c->gr[g].inv_perm.p : nullptr;
    }
    if (!m.pos4[1]) { m.pos4[1] = m.pos4[0]; m.pos4_ref[1] = m.pos4_ref[0]; m.inv_perm[1] = m.inv_perm[0]; }
    m.dmax = c->dmax.p; m.nsys = c->nsys;
    m.sysstate = c->sysstate.p; m.part = c->part_mv.p; m.step = c->step.p; m.nitems_total = c->n_items; m.mode = mode;
    launch_move(c->st, c->item_nwg, c->sys.p, c->wg_items.p, m);
    if (ev) (void)hipEventRecord((*ev)[3], c->st);
    if (mode == 0 && c->gr[1].present && c->hs[0].hybrid)
        launch_vs_place(c->st, c->item_nwg, c->sys.p, c->wg_items.p, c->s.p, c->n_items, c->label.p, c->itype.p,
                        c->vs_center.p, c->vs_moff.p, c->vs_members.p, c->mass.p, c->vs_invmass.p, c->vs_com.p,
                        c->gr[1].pos4.p,
                        c->gr[1].pos4_ref.p, c->gr[1].inv_perm.p, c->dmax.p + c->nsys);
    c->fin = fin_launch(c, mode);
    c->fin_pending = true;
    // NPT: the field sums decide the box rescaling, which must land before the next step's forces
    const bool npt = mode == 0 && c->npt_on;
    if (!defer || npt) flush_fin(c);
    if (npt)
        launch_npt(c->st, c->item_nwg, c->sys.p, c->wg_items.p, c->nsys, c->npt.p, c->s.p, c->vs_com.p, c->itype.p,
                   c->label.p, c->n_items, c->nb_tabs.p, c->dmax.p);
}

int check_flags(ibx_ctx *c) {
    int f[8];
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, memcpy_st(c, f, c->flags.p, sizeof f, hipMemcpyDeviceToHost));
    if (f[0] > 0)
        return fail(c, IBX_E_LIST_OVERFLOW, "LIST TOO SMALL: %d neighbours exceed the list capacity", f[0]);
    if (f[1] & 1) return fail(c, IBX_E_TABLE, "torsion angle outside its table");
    if (f[1] & 2) return fail(c, IBX_E_TABLE, "out-of-plane angle outside its table");
    return 0;
}

// Builds the lists; grows the capacity and rebuilds when a list overflowed.
int rebuild_checked(ibx_ctx *c) {
    for (int attempt = 0; attempt < 3; ++attempt) {
        HIPCHK(c, hipMemsetAsync(c->flags.p, 0, c->flags.bytes(), c->st));
        enqueue_rebuild(c);
        int f0 = 0;
        HIPCHK(c, hipStreamSynchronize(c->st));
        HIPCHK(c, memcpy_st(c, &f0, c->flags.p, sizeof f0, hipMemcpyDeviceToHost));
        if (f0 == 0) return 0;
        for (int g = 0; g < 2; ++g) {
            GroupDev &G = c->gr[g];
            if (!G.present || f0 <= G.cap) continue;
            G.cap = pad_to((int)(f0 * 1.25) + 16, 8);
            HIPCHK(c, G.list.alloc((size_t)G.nslot_pad * G.cap));
        }
        drop_graph(c);
    }
    return fail(c, IBX_E_LIST_OVERFLOW, "LIST TOO SMALL after resizing");
}

// One step.  With timing events each kernel class has its own launches, in
// order, so each event pair brackets one class.  Otherwise the force phase is
// one launch (group-0 nonbonded, bonded terms, and the previous step's field
// sums, which the previous step left pending), then the bead group's
// nonbonded force of a hybrid system, then the integrator; every launch
// follows the one before on the context's stream.
void enqueue_step(ibx_ctx *c, int step, std::vector<hipEvent_t> *ev) {
    // multiple time step: the nonbonded force on steps 1, k+1, 2k+1, ..., applied with weight k
    const bool nb_step = c->respa == 1 || (step - 1) % c->respa == 0;
    const float w = nb_step ? (float)c->respa : 0.0f;
    if (ev) {
        flush_fin(c);
        (void)hipEventRecord((*ev)[0], c->st);
        if (step % c->nupdate == 0) enqueue_rebuild(c, true);
        (void)hipEventRecord((*ev)[4], c->st);
        if (nb_step) enqueue_forces(c, 0, step);
        (void)hipEventRecord((*ev)[1], c->st);
        enqueue_terms(c);
        (void)hipEventRecord((*ev)[2], c->st);
        enqueue_rest(c, 0, ev, w);
        (void)hipEventRecord((*ev)[5], c->st);
        return;
    }
    if (step % c->nupdate == 0) enqueue_rebuild(c, true);
    if (nb_step) c->nb_par ^= 1;
    c->bd_par ^= 1;
    NbArgs a0;
    const bool g0 = nb_step && c->gr[0].present;
    if (g0) a0 = nb_args(c, 0);
    FinLaunch f = c->fin;
    const bool fin = c->fin_pending;
    c->fin_pending = false;
    launch_step_forces(c->st, g0 ? &a0 : nullptr, c->term_nwg, c->wg_terms.p, term_launch(c), c->s.p, c->n_items,
                       c->nterms, c->contrib.p, c->ncontrib, c->part_bd_buf(c->bd_par), c->flags.p,
                       fin ? &f : nullptr, c->ftiles);
    if (nb_step) enqueue_nb_groups(c, 0, step, 1);
    enqueue_rest(c, 0, nullptr, w, true);
}

// The force launch layouts the tile autotuner times: one workgroup per block, and persistent grids of
// 5 or 4 workgroups per CU taking one or two blocks at a time.
ForceTiles tile_candidate(const ibx_ctx *c, int k) {
    ForceTiles t;
    t.queue = c->fqueue.p;
    if (k == 0) return t;
    t.persistent = true;
    t.chunk = k == 2 ? 2 : 1;
    t.grid = c->ncu * (k == 3 ? 4 : 5);
    return t;
}

int ensure_events(ibx_ctx *c, size_t n);

// Tile autotuner (no counterpart in the reference).  Times each candidate
// layout of the force launch over a few launches of its force roles alone
// (nonbonded and bonded, no field sums) at the current positions and keeps the
// fastest.  Every layout computes the same bits, and these launches only
// overwrite what the next force evaluation overwrites again: the forces, the
// bonded contributions and the partial buffers that evaluation will write.
int tune_tiles(ibx_ctx *c) {
    c->tiles_pending = false;
    if (c->tiles_req[0] >= 0) {
        c->ftiles = tile_candidate(c, 0);
        if (c->tiles_req[0] == 1) {
            c->ftiles.persistent = true;
            c->ftiles.chunk = c->tiles_req[1] > 0 ? c->tiles_req[1] : 1;
            c->ftiles.grid = c->ncu * (c->tiles_req[2] > 0 ? c->tiles_req[2] : 5);
        }
        return 0;
    }
    int rc = ensure_events(c, 2);
    if (rc) return rc;
    const bool g0 = c->gr[0].present;
    NbArgs a0;
    if (g0) {
        a0 = nb_args(c, 0);
        a0.part = c->gr[0].part_buf(c->nb_par ^ 1);
    }
    constexpr int kReps = 3;
    int best = 0;
    for (int k = 0; k < kTileCandidates; ++k) {
        const ForceTiles t = tile_candidate(c, k);
        auto go = [&]() {
            launch_step_forces(c->st, g0 ? &a0 : nullptr, c->term_nwg, c->wg_terms.p, term_launch(c), c->s.p,
                               c->n_items, c->nterms, c->contrib.p, c->ncontrib, c->part_bd_buf(c->bd_par ^ 1),
                               c->flags.p, nullptr, t);
        };
        go();   // first touch
        HIPCHK(c, hipEventRecord(c->evpool[0], c->st));
        for (int r = 0; r < kReps; ++r) go();
        HIPCHK(c, hipEventRecord(c->evpool[1], c->st));
        HIPCHK(c, hipEventSynchronize(c->evpool[1]));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->evpool[0], c->evpool[1]));
        c->tile_ms[k] = ms / kReps;
        if (c->tile_ms[k] < c->tile_ms[best]) best = k;
    }
    c->ftiles = tile_candidate(c, best);
    return 0;
}

int ensure_events(ibx_ctx *c, size_t n) {
    while (c->evpool.size() < n) {
        hipEvent_t e;
        {
            std::lock_guard<std::mutex> lk(g_capture_mu);
            HIPCHK(c, hipEventCreate(&e));
        }
        c->evpool.push_back(e);
    }
    return 0;
}

}  // namespace

int ibx::device_free_bytes(size_t *free_bytes) {
    size_t fr = 0, total = 0;
    if (hipMemGetInfo(&fr, &total) != hipSuccess) return IBX_E_DEVICE;
    *free_bytes = fr;
    return IBX_OK;
}

// Context of the host-side systems hs (moved in): device layout, first lists.
int ibx::create_context(std::vector<HostSystem> &&hs, ibx_ctx **out) {
    *out = nullptr;
    const int nsys = (int)hs.size();
    if (nsys < 1) return fail(nullptr, IBX_E_INVALID, "no systems");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, IBX_E_DEVICE, "no HIP device");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return fail(nullptr, IBX_E_DEVICE, "no HIP device properties");
    if (!strstr(prop.gcnArchName, "gfx950"))
        return fail(nullptr, IBX_E_DEVICE, "device %s is not gfx950 (MI355X)", prop.gcnArchName);
    std::unique_ptr<ibx_ctx> c(new ibx_ctx());
    c->nsys = nsys;
    c->device = dev;
    c->ncu = prop.multiProcessorCount;
    c->hs = std::move(hs);
    for (int s = 0; s < nsys; ++s) {
        if (c->hs[s].nupdate != c->hs[0].nupdate)
            return fail(nullptr, IBX_E_INVALID, "all systems of a batch must share update_neighbour_list");
        if (c->hs[s].hybrid && nsys > 1) return fail(nullptr, IBX_E_INVALID, "hybrid systems run one per context");
        // VIRTUAL_DEF at set-up (main.F90:148-150)
        HostSystem &h = c->hs[s];
        for (int v = 0; v < h.nvs; ++v) {
            double p[3] = {0, 0, 0};
            if (h.vs_center[v] >= 0) {
                for (int k = 0; k < 3; ++k) p[k] = h.s[3 * (size_t)h.vs_center[v] + k];
            } else {
                for (int q = h.vs_off[v]; q < h.vs_off[v + 1]; ++q) {
                    int a = h.vs_atoms[q];
                    for (int k = 0; k < 3; ++k) p[k] += h.mass[h.itype[a]] * h.s[3 * (size_t)a + k];
                }
                for (int k = 0; k < 3; ++k) p[k] *= h.vs_invmass[v];
            }
            for (int k = 0; k < 3; ++k) h.s[3 * (size_t)(h.natoms + v) + k] = p[k];
        }
    }
    c->nupdate = c->hs[0].nupdate;
    std::unique_lock<std::mutex> lk(g_capture_mu);
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess)
        return fail(nullptr, IBX_E_DEVICE, "stream creation failed");
    lk.unlock();
    int rc = build_device(c.get());
    if (rc) { g_create_err = c->err; return rc; }
    if (nsys == 1 && c->hs[0].hybrid && c->hs[0].nmts > 0 && c->gr[1].present) {
        // MTS_LOOP.f90:20-22 buffers (zeroed: the first cycle's explicit steps fill them)
        ibx_ctx *cc = c.get();
        const size_t npart = (size_t)cc->gr[1].nwg * (kBlock / kNbAtomsPerBlock);
        if (cc->mts_f.alloc(9 * (size_t)cc->n_items) != hipSuccess || cc->mts_part.alloc(27 * npart) != hipSuccess ||
            memset_st(cc, cc->mts_f.p, 0, cc->mts_f.bytes()) != hipSuccess ||
            memset_st(cc, cc->mts_part.p, 0, cc->mts_part.bytes()) != hipSuccess)
            return fail(nullptr, IBX_E_NOMEM, "MTS buffers");
        cc->mts = true;
    }
    rc = rebuild_checked(c.get());
    if (rc) { g_create_err = c->err; return rc; }
    *out = c.release();
    return 0;
}

extern "C" {

const char *ibx_last_error(const ibx_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int ibx_create(int nsys, const ibx_system_desc *const *desc, ibx_ctx **out) {
    if (!out || nsys < 1 || !desc) return fail(nullptr, IBX_E_INVALID, "invalid arguments");
    *out = nullptr;
    std::vector<HostSystem> hs(nsys);
    for (int s = 0; s < nsys; ++s) {
        std::string err;
        int rc = build_host_system(desc[s], hs[s], err);
        if (rc) return fail(nullptr, rc, "system %d: %s", s, err.c_str());
    }
    return create_context(std::move(hs), out);
}

void ibx_destroy(ibx_ctx *ctx) {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->st);
    delete ctx;
}

int ibx_rebuild_lists(ibx_ctx *c) {
    if (!c) return fail(nullptr, IBX_E_INVALID, "null context");
    return rebuild_checked(c);
}

int ibx_compute_forces(ibx_ctx *c, int rebuild) {
    if (!c) return fail(nullptr, IBX_E_INVALID, "null context");
    if (rebuild) {
        int rc = rebuild_checked(c);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemsetAsync(c->flags.p, 0, c->flags.bytes(), c->st));
    enqueue_forces(c, 1);
    enqueue_terms(c);
    enqueue_rest(c, 1, nullptr);
    return check_flags(c);
}

int ibx_step(ibx_ctx *c, int nsteps, int step0) {
    if (!c || nsteps < 0) return fail(c, IBX_E_INVALID, "invalid arguments");
    if (c->tiles_pending && nsteps > 0) {
        int rc = tune_tiles(c);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemsetAsync(c->flags.p, 0, c->flags.bytes(), c->st));
    HIPCHK(c, hipMemcpyAsync(c->step.p, &step0, sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::vector<hipEvent_t> ev;
    const int per = 6;
    if (c->timing) {
        int rc = ensure_events(c, (size_t)per * nsteps);
        if (rc) return rc;
    }
    for (int k = 0; k < nsteps; ++k) {
        std::vector<hipEvent_t> e;
        if (c->timing) e.assign(c->evpool.begin() + per * k, c->evpool.begin() + per * (k + 1));
        enqueue_step(c, step0 + 1 + k, c->timing ? &e : nullptr);
    }
    flush_fin(c);
    c->counters[0] += nsteps;
    int rc = check_flags(c);
    if (c->timing) {
        for (int k = 0; k < nsteps; ++k) {
            hipEvent_t *e = c->evpool.data() + per * k;
            float t;
            if (hipEventElapsedTime(&t, e[4], e[1]) == hipSuccess) c->tm.ms[0] += t;
            if (hipEventElapsedTime(&t, e[1], e[2]) == hipSuccess) c->tm.ms[1] += t;
            if (hipEventElapsedTime(&t, e[2], e[3]) == hipSuccess) c->tm.ms[2] += t;
            if (hipEventElapsedTime(&t, e[0], e[4]) == hipSuccess) c->tm.ms[3] += t;
            if (hipEventElapsedTime(&t, e[0], e[5]) == hipSuccess) c->tm.ms[4] += t;
            c->tm.ms[5] += 1;
            if ((step0 + 1 + k) % c->nupdate == 0) c->tm.ms[7] += 1;
        }
        c->tm.ms[6] += nsteps;
    }
    return rc;
}

}  // extern "C"

// ---- scratch of ibx_step_noalloc -----------------------------------------------------------------
// The per-workgroup reduction partials (nonbonded of both groups, bonded, integrator), the per-system
// field sums and the per-system counters of finished field-sum workgroups live, for ibx_step_noalloc,
// in this array instead of the context's buffers.  Each device holds its own copy until the process
// ends; host code never passes the array's name to a kernel or a copy, only its device address, taken
// once per device with hipGetSymbolAddress.
static constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;                     // one ibx_step_noalloc at a time in the process
void *g_scratch_ptr[kMaxDevices] = {};       // device address of `scratch` on each device

int scratch_address(ibx_ctx *c, unsigned char **ptr) {
    int dev = 0;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev != c->device)
        return fail(c, IBX_E_INVALID, "the current device (%d) is not the context's (%d)", dev, c->device);
    if (dev < 0 || dev >= kMaxDevices) return fail(c, IBX_E_DEVICE, "device index %d out of range", dev);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *ptr = (unsigned char *)g_scratch_ptr[dev];
    return 0;
}

// The context buffers ibx_step_noalloc redirects into the scratch array, in order, with their
// offsets (each aligned to 256 bytes).  While a redirection is alive the buffers point into the
// array; its destructor points them back.
struct ScratchRedirect {
    struct Entry {
        std::function<void(void *)> set;   // stores an address into the buffer's pointer
        void *own;                         // the context's own allocation
        size_t bytes, off;
    };
    std::vector<Entry> e;
    size_t total = 0;
    template <typename T>
    void add(DBuf<T> &b) {
        if (!b.p) return;
        DBuf<T> *q = &b;
        e.push_back(Entry{[q](void *a) { q->p = (T *)a; }, (void *)b.p, b.bytes(), total});
        total += (b.bytes() + 255) / 256 * 256;
    }
    void point_to(unsigned char *base) {
        for (auto &x : e) x.set(base + x.off);
    }
    ~ScratchRedirect() {
        for (auto &x : e) x.set(x.own);
    }
};

}  // namespace

extern "C" {

int ibx_step_noalloc(ibx_ctx *c, int nsteps, int step0) {
    if (!c || nsteps < 0) return fail(c, IBX_E_INVALID, "invalid arguments");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    unsigned char *base = nullptr;
    int rc = scratch_address(c, &base);
    if (rc) return rc;
    flush_fin(c);   // nothing may be left reading the context's own partials
    ScratchRedirect r;
    for (int g = 0; g < 2; ++g)
        if (c->gr[g].present) r.add(c->gr[g].part);
    r.add(c->part_bd);
    r.add(c->part_mv);
    r.add(c->fsum);
    if (r.total > kScratchBytes)
        return fail(c, IBX_E_INVALID,
                    "ibx_step_noalloc: this context needs %zu bytes of scratch, more than the %zu of the static array",
                    r.total, kScratchBytes);
    // the array keeps the previous call's contents (maybe of another context): start from this context's
    // partials (the last evaluation's, which a step without a nonbonded evaluation still reads) and counters
    for (auto &x : r.e)
        HIPCHK(c, hipMemcpyAsync(base + x.off, x.own, x.bytes, hipMemcpyDeviceToDevice, c->st));
    r.point_to(base);
    launch_fill_int(c->st, c->flags.p, (int)c->flags.n, 0);
    HIPCHK(c, hipMemcpyAsync(c->step.p, &step0, sizeof(int), hipMemcpyHostToDevice, c->st));
    for (int k = 0; k < nsteps; ++k) enqueue_step(c, step0 + 1 + k, nullptr);
    flush_fin(c);
    for (auto &x : r.e)   // and hand the partials and field sums back to the context
        HIPCHK(c, hipMemcpyAsync(x.own, base + x.off, x.bytes, hipMemcpyDeviceToDevice, c->st));
    c->counters[0] += nsteps;
    return check_flags(c);   // waits for the stream, so no two calls overlap on the array
}

// Steps per long graph block (each block ends with its own field sums).
constexpr int kGraphSteps = 40;

int long_block(const ibx_ctx *c) {
    const int period = block_period(c);
    return period * std::max(1, kGraphSteps / period);
}

// nblk replays of a graph of `steps` steps (a whole number of block periods)
// starting after step `cur`.  The step schedule (rebuilds, multiple-time-step
// cycles) repeats with the block period, so one graph serves every start step
// of the same phase cur % period; it is captured on first use, or again when
// the partial buffers it expects have changed.
int run_graph_blocks(ibx_ctx *c, int steps, int nblk, int cur) {
    const int phase = cur % block_period(c);
    ibx_ctx::GraphSlot *g = nullptr;
    for (auto &q : c->graph)
        if (q.exec && q.steps == steps && q.phase == phase && q.par == c->nb_par && q.bd_par == c->bd_par) g = &q;
    if (!g) {
        // an empty slot, else the least recently used one
        g = &c->graph[0];
        for (auto &q : c->graph)
            if (!q.exec || (g->exec && q.used < g->used)) g = &q;
        std::lock_guard<std::mutex> lk(g_capture_mu);
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        *g = ibx_ctx::GraphSlot();
        hipGraph_t graph;
        const int par = c->nb_par, bd_par = c->bd_par;
        // relaxed: other host threads (other contexts, the serving loop) keep allocating and copying
        // while this one captures; nothing captured here is a call that capture forbids
        HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed));
        for (int k = 0; k < steps; ++k) enqueue_step(c, phase + 1 + k, nullptr);
        flush_fin(c);
        HIPCHK(c, hipStreamEndCapture(c->st, &graph));
        hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g->exec = nullptr;
            return fail(c, IBX_E_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
        }
        // a block holds an even number of evaluations, so it ends on the buffers it started from
        if (c->nb_par != par || c->bd_par != bd_par)
            return fail(c, IBX_E_DEVICE, "graph block of %d steps does not return to its partial buffers", steps);
        g->steps = steps;
        g->phase = phase;
        g->par = par;
        g->bd_par = bd_par;
        ++c->counters[3];
    }
    g->used = ++c->graph_clock;
    HIPCHK(c, hipMemsetAsync(c->flags.p, 0, c->flags.bytes(), c->st));
    HIPCHK(c, hipMemcpyAsync(c->step.p, &cur, sizeof(int), hipMemcpyHostToDevice, c->st));
    for (int b = 0; b < nblk; ++b) HIPCHK(c, hipGraphLaunch(g->exec, c->st));
    c->counters[1] += (int64_t)nblk * steps;
    c->counters[2] += nblk;
    return check_flags(c);
}

int ibx_step_graph(ibx_ctx *c, int nsteps, int step0) {
    if (!c || nsteps < 0 || step0 < 0) return fail(c, IBX_E_INVALID, "invalid arguments");
    if (c->tiles_pending && nsteps > 0) {   // before any capture
        int rc = tune_tiles(c);
        if (rc) return rc;
    }
    // graph blocks of several block periods, then of one period, from step0 whatever its phase; eager tail
    const int period = block_period(c);
    const int longer = long_block(c);
    int done = 0, rc;
    for (int steps : {longer, period}) {
        const int nblk = (nsteps - done) / steps;
        if (nblk == 0) continue;
        if ((rc = run_graph_blocks(c, steps, nblk, step0 + done))) return rc;
        done += nblk * steps;
    }
    return ibx_step(c, nsteps - done, step0 + done);
}

int ibx_get_counters(ibx_ctx *c, int64_t out[8]) {
    if (!c || !out) return fail(c, IBX_E_INVALID, "invalid arguments");
    for (int k = 0; k < 4; ++k) out[k] = c->counters[k];
    out[4] = block_period(c);
    out[5] = long_block(c);
    out[6] = 0;
    out[7] = 0;
    return 0;
}

int ibx_set_respa(ibx_ctx *c, int k) {
    if (!c || k < 1 || k > 64) return fail(c, IBX_E_INVALID, "invalid multiple-time-step ratio");
    if (c->mts && k > 1) return fail(c, IBX_E_INVALID, "the MTS cycle of a hybrid system excludes RESPA");
    c->respa = k;
    drop_graph(c);
    return 0;
}

int ibx_set_skin_tuning(ibx_ctx *c, int on) {
    if (!c) return fail(nullptr, IBX_E_INVALID, "null context");
    c->skin_tune = on != 0;
    drop_graph(c);
    return 0;
}

int ibx_set_tiles(ibx_ctx *c, const int tiles[IBX_NTILES]) {
    if (!c || !tiles) return fail(c, IBX_E_INVALID, "null argument");
    if (tiles[0] < -1 || tiles[0] > 1) return fail(c, IBX_E_INVALID, "force launch layout %d is not -1, 0 or 1", tiles[0]);
    if (tiles[1] < -1 || tiles[1] == 0 || tiles[1] > 64) return fail(c, IBX_E_INVALID, "blocks per take %d", tiles[1]);
    if (tiles[2] < -1 || tiles[2] == 0 || tiles[2] > 16) return fail(c, IBX_E_INVALID, "workgroups per CU %d", tiles[2]);
    for (int k = 0; k < IBX_NTILES; ++k) c->tiles_req[k] = tiles[k];
    c->tiles_pending = true;   // applied (or tuned) at the next step call
    for (float &m : c->tile_ms) m = -1.f;
    drop_graph(c);
    return 0;
}

int ibx_get_tiles(ibx_ctx *c, int tiles[IBX_NTILES], float ms[IBX_TILE_CANDIDATES]) {
    if (!c || !tiles) return fail(c, IBX_E_INVALID, "null argument");
    if (c->tiles_pending) {
        for (int k = 0; k < IBX_NTILES; ++k) tiles[k] = -1;
    } else {
        tiles[0] = c->ftiles.persistent ? 1 : 0;
        tiles[1] = c->ftiles.persistent ? c->ftiles.chunk : 0;
        tiles[2] = c->ftiles.persistent && c->ncu > 0 ? c->ftiles.grid / c->ncu : 0;
    }
    if (ms)
        for (int k = 0; k < IBX_TILE_CANDIDATES; ++k) ms[k] = c->tile_ms[k];
    return 0;
}

int ibx_get_skin(ibx_ctx *c, int sys, double delta[2]) {
    if (!c || sys < 0 || sys >= c->nsys || !delta) return fail(c, IBX_E_INVALID, "invalid system index");
    HIPCHK(c, hipStreamSynchronize(c->st));
    DevSys D;
    HIPCHK(c, memcpy_st(c, &D, c->sys.p + sys, sizeof D, hipMemcpyDeviceToHost));
    for (int g = 0; g < 2; ++g)
        delta[g] = g < D.ngroup ? std::sqrt((double)D.g[g].rin2u) * D.unit - std::sqrt(D.g[g].rc2) : 0.0;
    return 0;
}

int ibx_retire(ibx_ctx *c, int sys) {
    if (!c || sys < 0 || sys >= c->nsys) return fail(c, IBX_E_INVALID, "invalid system index");
    if (!c->dsys[sys].live) return 0;
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->dsys[sys].live = 0;
    const int zero = 0;
    HIPCHK(c, memcpy_st(c, (char *)(c->sys.p + sys) + offsetof(DevSys, live), &zero, sizeof zero, hipMemcpyHostToDevice));
    // the integrator's and the bonded terms' workgroups of the system return before loading anything
    const DevSys &D = c->dsys[sys];
    launch_fill_int(c->st, c->wg_items.p + D.item_wg_off, D.item_nwg, sys | kWgRetired);
    launch_fill_int(c->st, c->wg_terms.p + D.term_wg_off, D.term_nwg, sys | kWgRetired);
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

}  // extern "C"

namespace {

// Host copy of n elements into b at element offset off (the context's stream, then a wait).
template <typename T>
hipError_t put(ibx_ctx *c, DBuf<T> &b, size_t off, const std::vector<T> &h) {
    if (h.empty()) return hipSuccess;
    if (off + h.size() > b.n) return hipErrorInvalidValue;
    return hipMemcpyAsync(b.p + off, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->st);
}

// n ints of value v from p on (the fill kernel; also zeroes floats and doubles).
void fill_words(ibx_ctx *c, void *p, size_t nwords, int v) { launch_fill_int(c->st, (int *)p, (int)nwords, v); }

// The device's record of system sys into c->dsys[sys]: an NPT system's box, and the constants that follow it,
// change on the device.  The caller has synchronised the stream.
int pull_sys(ibx_ctx *c, int sys) {
    HIPCHK(c, memcpy_st(c, &c->dsys[sys], c->sys.p + sys, sizeof(DevSys), hipMemcpyDeviceToHost));
    return 0;
}

}  // namespace

// Place k of a batch taken over by system h (see ibx_assign).
int ibx::assign_system(ibx_ctx *c, int k, HostSystem &&h, int step0) {
    if (!c || k < 0 || k >= c->nsys) return fail(c, IBX_E_INVALID, "invalid system index");
    if (c->dsys[k].live) return fail(c, IBX_E_INVALID, "place %d still holds a live system (retire it first)", k);
    if (h.nvs > 0 || h.hybrid || h.ngroup != 1)
        return fail(c, IBX_E_INVALID, "only systems without virtual sites and with one nonbonded group can take over a place");
    if (c->mts || c->gr[1].present || !c->gr[0].present)
        return fail(c, IBX_E_INVALID, "places of this batch cannot be taken over");
    if (h.nupdate != c->nupdate) return fail(c, IBX_E_INVALID, "update_neighbour_list differs from the batch's");
    if (step0 < 0 || step0 % c->nupdate || step0 % c->respa)
        return fail(c, IBX_E_INVALID, "step0 must be a multiple of the list interval and of the RESPA factor");
    const SysLayout &L = c->lay[k];
    const SlotSizes need = slot_needs(h);
    if (!slot_fits(L.cap, need) || h.nitems >= L.cap.items)
        return fail(c, IBX_E_NOMEM, "the system does not fit place %d", k);
    GroupDev &G = c->gr[0];
    if (list_capacity_needed(h) > G.cap)
        return fail(c, IBX_E_NOMEM, "the system's lists need more than the batch's list capacity %d", G.cap);
    if (G.max_rows > 0 && need.nb_rows > G.max_rows)
        return fail(c, IBX_E_NOMEM, "nonbonded tables of %d rows exceed the batch's %d", need.nb_rows, G.max_rows);
    if (need.nb_tabs > 1 && !G.multi) return fail(c, IBX_E_INVALID, "several nonbonded tables in a batch of one each");
    if (need.nb_tabs > kMaxNbTabs) return fail(c, IBX_E_INVALID, "more than %d nonbonded tables", kMaxNbTabs);
    if (need.nb_tabs > 1 && h.ntype > kMaxNbTypes)
        return fail(c, IBX_E_INVALID, "more than %d types with several nonbonded tables", kMaxNbTypes);
    HIPCHK(c, hipStreamSynchronize(c->st));

    const int o = L.item_off, np = L.cap.items, N = c->n_items;
    // items
    std::vector<double> hs(3 * (size_t)np, 0.0), hv(3 * (size_t)np, 0.0), hm(np, 1.0), him(np, 1.0), hzero(np, 0.0);
    std::vector<int> hit(np, -1), hlab(np, 0), hneg(np, -1), hvm(np, L.vsm_off), hexo(np), hinco(np);
    for (int i = 0; i < h.nitems; ++i) {
        for (int q = 0; q < 3; ++q) hs[(size_t)q * np + i] = h.s[3 * (size_t)i + q];
        if (i < h.natoms)
            for (int q = 0; q < 3; ++q) hv[(size_t)q * np + i] = h.v[3 * (size_t)i + q];
        hit[i] = h.itype[i];
        hlab[i] = h.type_label[i];
        hm[i] = h.mass[h.itype[i]];
        him[i] = h.invmass[h.itype[i]];
    }
    // exclusions in global indices; the room left belongs to the last (padding) item
    std::vector<int> hex;
    for (int i = 0; i < np; ++i) {
        hexo[i] = L.excl_off + (int)hex.size();
        if (i < h.nitems)
            for (int q = h.excl_off[i]; q < h.excl_off[i + 1]; ++q) hex.push_back(o + h.excl[q]);
    }
    hex.resize(L.cap.excl, o);
    // bonded tables and terms, members' contribution slots contiguous per item in term order
    std::vector<BdTab> hbt;
    std::vector<float4> hbrows;
    std::vector<double> hbx;
    auto add_tabs = [&](const std::vector<Table> &tabs, const std::vector<int> &labs) {
        const int base = L.bd_tab_off + (int)hbt.size();
        for (size_t t = 0; t < tabs.size(); ++t) {
            BdTab b;
            b.row_off = L.bd_row_off + (int)hbrows.size();
            b.ndat = tabs[t].ndat;
            b.x0 = tabs[t].x0;
            b.bin = tabs[t].bin;
            b.inv_bin = 1.0 / tabs[t].bin;
            b.pad = 0;
            b.label = labs.empty() ? 1 : labs[t];
            hbt.push_back(b);
            append_rows(hbrows, tabs[t]);
            for (int q = 0; q <= tabs[t].ndat; ++q) hbx.push_back(tabs[t].x[q]);
        }
        return base;
    };
    const int bbase = add_tabs(h.bt, h.bond_lab), abase = add_tabs(h.at, h.angle_lab);
    const int tbase = add_tabs(h.tt, h.tors_lab), obase = add_tabs(h.ot, h.oop_lab);
    if (L.bd_tab_off + L.cap.bd_tabs - 1 > kTermTabMask) return fail(c, IBX_E_INVALID, "too many bonded tables");
    hbt.resize(L.cap.bd_tabs, BdTab{});
    hbrows.resize(L.cap.bd_rows, make_float4(0.f, 0.f, 0.f, 0.f));
    hbx.resize(L.cap.bd_rows, 0.0);
    std::vector<int4> htm;
    std::vector<int> hmeta, hslot;
    std::vector<std::vector<int>> inc(np);
    int nslot = 0;
    auto add_term = [&](int kind, int tab, const std::array<int, 4> &m, int nm) {
        htm.push_back(make_int4(o + m[0], o + m[1], nm > 2 ? o + m[2] : -1, nm > 3 ? o + m[3] : -1));
        hmeta.push_back(kind | (tab << 2));
        hslot.push_back(nslot);
        for (int r = 0; r < nm; ++r) inc[m[r]].push_back(nslot + r);
        nslot += nm;
    };
    for (size_t q = 0; q < h.bonds.size(); ++q) add_term(KBOND, bbase + h.bond_tab[q], h.bonds[q], 2);
    for (size_t q = 0; q < h.angles.size(); ++q) add_term(KANGLE, abase + h.angle_tab[q], h.angles[q], 3);
    for (size_t q = 0; q < h.tors.size(); ++q) add_term(KTORS, tbase + h.tors_tab[q], h.tors[q], 4);
    for (size_t q = 0; q < h.oops.size(); ++q) add_term(KOOP, obase + h.oop_tab[q], h.oops[q], 4);
    std::vector<int> renum(std::max(1, nslot), 0);
    for (int i = 0, at = 0; i < np; ++i) {
        hinco[i] = L.contrib_off + at;
        for (size_t q = 0; q < inc[i].size(); ++q) renum[inc[i][q]] = at + (int)q;
        at += (int)inc[i].size();
    }
    std::vector<int4> hslot4(L.cap.terms, make_int4(0, 0, 0, 0));
    for (size_t t = 0; t < hslot.size(); ++t) {
        const int4 m = htm[t];
        const int b = hslot[t], co = L.contrib_off;
        hslot4[t] = make_int4(co + renum[b], co + renum[b + 1], m.z >= 0 ? co + renum[b + 2] : 0,
                              m.w >= 0 ? co + renum[b + 3] : 0);
    }
    htm.resize(L.cap.terms, make_int4(-1, -1, -1, -1));
    hmeta.resize(L.cap.terms, -1);
    // nonbonded tables
    DevSys D{};
    set_sys_constants(h, D);
    std::vector<float4> hnrows;
    std::vector<float> hnxoff;
    std::vector<NbTab> hntab;
    for (const Table &t : h.nt) {
        hntab.push_back(nb_header(t, D.unit, L.nb_row_off + (int)hnrows.size(), hnxoff));
        append_nb_rows(hnrows, t);
    }
    D.nb_row_off = L.nb_row_off;
    D.nb_nrow = (int)hnrows.size();
    D.nb_tab_off = L.nb_tab_off;
    D.nb_ntab = (int)h.nt.size();
    D.ntype = h.ntype;
    D.nbmap_off = L.nb_map_off;
    hnrows.resize(L.cap.nb_rows, make_float4(0.f, 0.f, 0.f, 0.f));
    hnxoff.resize(L.cap.nb_rows, 0.f);
    hntab.resize(L.cap.nb_tabs, NbTab{});
    std::vector<int> hnmap(L.cap.nb_map, -1);
    for (int q = 0; q < h.ntype * h.ntype; ++q) hnmap[q] = h.inbondt[q] < 0 ? -1 : L.nb_tab_off + h.inbondt[q];
    // the record: offsets of the place, the system's constants
    const DevSys &P = c->dsys[k];
    D.item_off = o;
    D.nitems_pad = np;
    D.item_wg_off = P.item_wg_off;
    D.item_nwg = P.item_nwg;
    D.term_wg_off = P.term_wg_off;
    D.term_nwg = P.term_nwg;
    for (int g = 0; g < 2; ++g) {
        DevGroup &dg = D.g[g];
        dg = P.g[g];   // offsets of the place
        if (g < h.ngroup) set_group_constants(h, g, D.unit, dg);
        dg.nmem = need.members[g];
        dg.cell_off = L.cell_off[g];
    }
    D.step0 = step0;
    D.step_ctr = c->step.p;
    std::vector<int> hmem(L.cap.members[0], -1);
    for (int q = 0; q < need.members[0]; ++q) hmem[q] = o + h.members[0][q];

    // device writes, all on the context's stream
    hipError_t e = hipSuccess;
    for (int q = 0; q < 3 && e == hipSuccess; ++q) {
        std::vector<double> a(hs.begin() + (size_t)q * np, hs.begin() + (size_t)(q + 1) * np);
        std::vector<double> b(hv.begin() + (size_t)q * np, hv.begin() + (size_t)(q + 1) * np);
        e = put(c, c->s, (size_t)q * N + o, a);
        if (e == hipSuccess) e = put(c, c->v, (size_t)q * N + o, b);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);   // a and b leave scope
    }
    auto chain = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    chain(put(c, c->mass, o, hm));
    chain(put(c, c->invmass, o, him));
    chain(put(c, c->itype, o, hit));
    chain(put(c, c->label, o, hlab));
    chain(put(c, c->vs_of, o, hneg));
    chain(put(c, c->vs_center, o, hneg));
    chain(put(c, c->vs_coeff, o, hzero));
    chain(put(c, c->vs_invmass, o, hzero));
    chain(put(c, c->vs_moff, o, hvm));
    chain(put(c, c->excl_off, o, hexo));
    chain(put(c, c->excl, L.excl_off, hex));
    chain(put(c, c->inc_off, o, hinco));
    chain(put(c, c->term_mem, L.term_off, htm));
    chain(put(c, c->term_meta, L.term_off, hmeta));
    chain(put(c, c->term_slot, L.term_off, hslot4));
    chain(put(c, c->bd_tabs, L.bd_tab_off, hbt));
    chain(put(c, c->bd_rows, L.bd_row_off, hbrows));
    chain(put(c, c->bd_x, L.bd_row_off, hbx));
    chain(put(c, c->nb_rows, L.nb_row_off, hnrows));
    chain(put(c, c->nb_xoff, L.nb_row_off, hnxoff));
    chain(put(c, c->nb_tabs, L.nb_tab_off, hntab));
    chain(put(c, c->nb_map, L.nb_map_off, hnmap));
    chain(put(c, G.mem, L.slot_off[0], hmem));
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(c->st);   // no copy left reading the host arrays
        return fail(c, IBX_E_DEVICE, "assign: %s", hipGetErrorString(e));
    }
    // state of the place left by the system before: cleared
    fill_words(c, c->contrib.p + L.contrib_off, sizeof(Contrib) / sizeof(int) * (size_t)L.cap.contrib, 0);
    for (int q = 0; q < 3; ++q) {
        fill_words(c, c->fnb.p + (size_t)q * N + o, np, 0);
        fill_words(c, c->ftot.p + (size_t)q * N + o, np, 0);
    }
    fill_words(c, G.perm.p + L.slot_off[0], L.cap.members[0], -1);
    fill_words(c, G.nneb.p + L.slot_off[0], 2 * (size_t)L.cap.members[0], 0);
    fill_words(c, G.inv_perm.p + o, np, -1);
    fill_words(c, c->obs.p + (size_t)k * kObs, 2 * (size_t)kObs, 0);
    fill_words(c, c->avg.p + (size_t)k * kObs, 2 * (size_t)kObs, 0);
    fill_words(c, c->npt.p + (size_t)k * kNptState, 2 * (size_t)kNptState, 0);
    if (c->rdf_nbin) {
        fill_words(c, c->rdf.p + (size_t)k * c->rdf_nbin, 2 * (size_t)c->rdf_nbin, 0);
        fill_words(c, c->rdf_samples.p + k, 1, 0);
        fill_words(c, c->rdf_ivol.p + k, 2, 0);
    }
    std::vector<double> st(8, 0.0);
    st[0] = h.temp_atom0;
    st[1] = h.temp_bead0;
    st[5] = 1.0;
    chain(put(c, c->sysstate, 8 * (size_t)k, st));
    std::vector<DevSys> one(1, D);
    chain(put(c, c->sys, (size_t)k, one));
    launch_fill_int(c->st, c->wg_items.p + D.item_wg_off, D.item_nwg, k);   // no longer retired
    launch_fill_int(c->st, c->wg_terms.p + D.term_wg_off, D.term_nwg, k);
    chain(hipStreamSynchronize(c->st));
    if (e != hipSuccess) return fail(c, IBX_E_DEVICE, "assign: %s", hipGetErrorString(e));
    for (int t = 0; t < (int)h.nt.size(); ++t) c->h_nb_tabs[L.nb_tab_off + t] = hntab[t];
    // a non-cubic box in a batch of cubic ones needs the general kernels: graphs are captured again
    bool cubic_before = true, cubic_after = true;
    for (int s = 0; s < c->nsys; ++s) {
        const DevSys &X = s == k ? D : c->dsys[s];
        cubic_before = cubic_before && c->dsys[s].rho[1] == 1.0f && c->dsys[s].rho[2] == 1.0f;
        cubic_after = cubic_after && X.rho[1] == 1.0f && X.rho[2] == 1.0f;
    }
    if (cubic_before != cubic_after) drop_graph(c);
    if (h.ensemble == 2 && !c->npt_on) {   // the steps gain the box rescaling: graphs are captured again
        c->npt_on = true;
        drop_graph(c);
    }
    c->dsys[k] = D;
    c->hs[k] = std::move(h);
    // the place's first lists; the other systems' lists are untouched
    HIPCHK(c, hipMemsetAsync(c->flags.p, 0, c->flags.bytes(), c->st));
    RebuildArgs a = rebuild_args(c, 0);
    a.only = k;
    launch_rebuild(c->st, a);
    int f0 = 0;
    HIPCHK(c, memcpy_st(c, &f0, c->flags.p, sizeof f0, hipMemcpyDeviceToHost));
    if (f0 > 0) {
        (void)ibx_retire(c, k);
        return fail(c, IBX_E_LIST_OVERFLOW, "LIST TOO SMALL: %d neighbours exceed the list capacity", f0);
    }
    return 0;
}

SlotSizes ibx::slot_capacity(const ibx_ctx *c, int k) { return c->lay[k].cap; }

const HostSystem *ibx::host_system(const ibx_ctx *c, int k) { return (c && k >= 0 && k < c->nsys) ? &c->hs[k] : nullptr; }

int ibx::batch_period(const ibx_ctx *c) { return block_period(c); }

extern "C" {

int ibx_assign(ibx_ctx *c, int sys, const ibx_system_desc *desc, int step0) {
    if (!c || !desc) return fail(c, IBX_E_INVALID, "invalid arguments");
    HostSystem h;
    std::string err;
    int rc = build_host_system(desc, h, err);
    if (rc) return fail(c, rc, "%s", err.c_str());
    return assign_system(c, sys, std::move(h), step0);
}

int ibx_get_capacity(ibx_ctx *c, int sys, int64_t cap[16]) {
    if (!c || !cap || sys < 0 || sys >= c->nsys) return fail(c, IBX_E_INVALID, "invalid system index");
    const SlotSizes &z = c->lay[sys].cap;
    const int64_t v[16] = {z.items, z.terms, z.contrib, z.excl, z.members[0], z.members[1], z.cells[0], z.cells[1],
                           z.nb_rows, z.nb_tabs, z.nb_map, z.bd_tabs, z.bd_rows, c->gr[0].cap, block_period(c), 0};
    for (int q = 0; q < 16; ++q) cap[q] = v[q];
    return 0;
}

int ibx_set_last_step(ibx_ctx *c, int sys, int nsteps) {
    if (!c || sys < 0 || sys >= c->nsys) return fail(c, IBX_E_INVALID, "invalid system index");
    HIPCHK(c, hipStreamSynchronize(c->st));
    DevSys &D = c->dsys[sys];
    D.last_step = nsteps < 0 ? INT_MAX : nsteps;
    int cur = 0;
    HIPCHK(c, memcpy_st(c, &cur, c->step.p, sizeof cur, hipMemcpyDeviceToHost));
    const double go = alive_at(D, cur + 1) ? 1.0 : 0.0;   // whether the next move advances it
    HIPCHK(c, memcpy_st(c, (char *)(c->sys.p + sys) + offsetof(DevSys, last_step), &D.last_step, sizeof(int),
                        hipMemcpyHostToDevice));
    HIPCHK(c, memcpy_st(c, c->sysstate.p + 8 * (size_t)sys + 5, &go, sizeof go, hipMemcpyHostToDevice));
    return 0;
}

int ibx_get_observables(ibx_ctx *c, int sys, double obs[32]) {
    if (!c || sys < 0 || sys >= c->nsys) return fail(c, IBX_E_INVALID, "invalid system index");
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, memcpy_st(c, obs, c->obs.p + (size_t)sys * kObs, kObs * sizeof(double), hipMemcpyDeviceToHost));
    double a[kObs];
    HIPCHK(c, memcpy_st(c, a, c->avg.p + (size_t)sys * kObs, kObs * sizeof(double), hipMemcpyDeviceToHost));
    if (int rc = pull_sys(c, sys)) return rc;
    obs[26] = c->dsys[sys].volume;
    obs[28] = a[28];
    return 0;
}

int ibx_get_box(ibx_ctx *c, int sys, double box[3]) {
    if (!c || sys < 0 || sys >= c->nsys || !box) return fail(c, IBX_E_INVALID, "invalid arguments");
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (int rc = pull_sys(c, sys)) return rc;
    for (int k = 0; k < 3; ++k) box[k] = c->dsys[sys].box[k];
    return 0;
}

int ibx_get_averages(ibx_ctx *c, int sys, double sums[32], int reset) {
    if (!c || sys < 0 || sys >= c->nsys) return fail(c, IBX_E_INVALID, "invalid system index");
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, memcpy_st(c, sums, c->avg.p + (size_t)sys * kObs, kObs * sizeof(double), hipMemcpyDeviceToHost));
    if (reset) HIPCHK(c, memset_st(c, c->avg.p + (size_t)sys * kObs, 0, kObs * sizeof(double)));
    return 0;
}

int ibx_get_forces(ibx_ctx *c, int sys, double *f) {
    if (!c || sys < 0 || sys >= c->nsys || !f) return fail(c, IBX_E_INVALID, "invalid arguments");
    HIPCHK(c, hipStreamSynchronize(c->st));
    const HostSystem &h = c->hs[sys];
    const int o = c->item_off[sys], N = c->n_items;
    std::vector<float> tmp(h.nitems);
    for (int k = 0; k < 3; ++k) {
        HIPCHK(c, memcpy_st(c, tmp.data(), c->ftot.p + (size_t)k * N + o, h.nitems * sizeof(float), hipMemcpyDeviceToHost));
        for (int i = 0; i < h.nitems; ++i) f[3 * (size_t)i + k] = tmp[i];
    }
    return 0;
}

int ibx_get_state(ibx_ctx *c, int sys, double *s, double *v) {
    if (!c || sys < 0 || sys >= c->nsys) return fail(c, IBX_E_INVALID, "invalid system index");
    HIPCHK(c, hipStreamSynchronize(c->st));
    const HostSystem &h = c->hs[sys];
    const int o = c->item_off[sys], N = c->n_items;
    std::vector<double> tmp(h.nitems);
    double st[8];
    HIPCHK(c, memcpy_st(c, st, c->sysstate.p + 8 * (size_t)sys, sizeof st, hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; ++k) {
        if (s) {
            HIPCHK(c, memcpy_st(c, tmp.data(), c->s.p + (size_t)k * N + o, h.nitems * sizeof(double), hipMemcpyDeviceToHost));
            for (int i = 0; i < h.nitems; ++i) s[3 * (size_t)i + k] = tmp[i];
        }
        if (v) {
            HIPCHK(c, memcpy_st(c, tmp.data(), c->v.p + (size_t)k * N + o, h.natoms * sizeof(double), hipMemcpyDeviceToHost));
            // a MOMENTUM correction decided at the last step is applied lazily by the next move
            for (int i = 0; i < h.natoms; ++i) v[3 * (size_t)i + k] = tmp[i] - st[2 + k] * h.invmass[h.itype[i]];
        }
    }
    return 0;
}

int ibx_set_state(ibx_ctx *c, int sys, const double *s, const double *v) {
    if (!c || sys < 0 || sys >= c->nsys) return fail(c, IBX_E_INVALID, "invalid system index");
    HIPCHK(c, hipStreamSynchronize(c->st));
    HostSystem &h = c->hs[sys];
    const int o = c->item_off[sys], N = c->n_items;
    std::vector<double> tmp(h.nitems);
    if (s) {
        for (int i = 0; i < h.natoms; ++i)
            for (int k = 0; k < 3; ++k) h.s[3 * (size_t)i + k] = s[3 * (size_t)i + k];
        for (int vv = 0; vv < h.nvs; ++vv) {
            double p[3] = {0, 0, 0};
            if (h.vs_center[vv] >= 0) {
                for (int k = 0; k < 3; ++k) p[k] = h.s[3 * (size_t)h.vs_center[vv] + k];
            } else {
                for (int q = h.vs_off[vv]; q < h.vs_off[vv + 1]; ++q) {
                    int a = h.vs_atoms[q];
                    for (int k = 0; k < 3; ++k) p[k] += h.mass[h.itype[a]] * h.s[3 * (size_t)a + k];
                }
                for (int k = 0; k < 3; ++k) p[k] *= h.vs_invmass[vv];
            }
            for (int k = 0; k < 3; ++k) h.s[3 * (size_t)(h.natoms + vv) + k] = p[k];
        }
        for (int k = 0; k < 3; ++k) {
            for (int i = 0; i < h.nitems; ++i) tmp[i] = h.s[3 * (size_t)i + k];
            HIPCHK(c, memcpy_st(c, c->s.p + (size_t)k * N + o, tmp.data(), h.nitems * sizeof(double), hipMemcpyHostToDevice));
        }
        if (h.nvs > 0) {
            std::vector<double> com(3 * (size_t)h.nvs);
            for (int vv = 0; vv < h.nvs; ++vv) vs_centre_of_atoms(h, vv, &com[3 * (size_t)vv]);
            for (int k = 0; k < 3; ++k) {
                for (int vv = 0; vv < h.nvs; ++vv) tmp[vv] = com[3 * (size_t)vv + k];
                HIPCHK(c, memcpy_st(c, c->vs_com.p + (size_t)k * N + o + h.natoms, tmp.data(), h.nvs * sizeof(double),
                                    hipMemcpyHostToDevice));
            }
        }
    }
    if (v) {
        for (int k = 0; k < 3; ++k) {
            for (int i = 0; i < h.natoms; ++i) tmp[i] = v[3 * (size_t)i + k];
            HIPCHK(c, memcpy_st(c, c->v.p + (size_t)k * N + o, tmp.data(), h.natoms * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    HIPCHK(c, memset_st(c, c->sysstate.p + 8 * (size_t)sys + 2, 0, 3 * sizeof(double)));
    if (s) return rebuild_checked(c);   // lists of the new positions (every system of the context)
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

int ibx_set_nb_table(ibx_ctx *c, int sys, int t, int nrow, const double *x, const double *v) {
    if (!c || sys < 0 || sys >= c->nsys) return fail(c, IBX_E_INVALID, "invalid system index");
    HostSystem &h = c->hs[sys];
    if (t < 0 || t >= (int)h.nt.size()) return fail(c, IBX_E_INVALID, "invalid table index");
    // same load + FTABLE as set-up
    Table nt;
    {
        nt.x.assign(nrow + 4, 0.0);
        nt.V.assign(nrow + 4, 0.0);
        nt.F.assign(nrow + 4, 0.0);
        double bin = x[1] - x[0];
        int k;
        if (x[0] != 0.0) {
            nt.x[0] = 0.0; nt.V[0] = v[0]; nt.x[1] = bin; nt.V[1] = v[0]; nt.x[2] = x[1]; nt.V[2] = v[1];
            for (k = 3; k - 1 < nrow; ++k) { nt.x[k] = x[k - 1]; nt.V[k] = v[k - 1]; }
        } else {
            for (k = 0; k < nrow; ++k) { nt.x[k] = x[k]; nt.V[k] = v[k]; }
        }
        nt.ndat = k - 1;
        nt.bin = bin;
        for (int j = 1; j <= nt.ndat; ++j) nt.F[j] = -(nt.V[j + 1] - nt.V[j - 1]) / nt.bin / nt.x[j] / 2.0;
        nt.F[0] = nt.F[1];
    }
    if (nt.ndat != h.nt[t].ndat) return fail(c, IBX_E_INVALID, "replacement table must keep the row count");
    h.nt[t] = nt;
    std::vector<float4> rows;
    std::vector<float> xoff;
    append_nb_rows(rows, nt);
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (int rc = pull_sys(c, sys)) return rc;   // the current fixed-point unit
    const NbTab &hdr = c->h_nb_tabs[c->dsys[sys].nb_tab_off + t];
    const NbTab nh = nb_header(nt, c->dsys[sys].unit, hdr.row_off, xoff);
    HIPCHK(c, memcpy_st(c, c->nb_rows.p + hdr.row_off, rows.data(), rows.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(c, memcpy_st(c, c->nb_xoff.p + hdr.row_off, xoff.data(), xoff.size() * sizeof(float), hipMemcpyHostToDevice));
    c->h_nb_tabs[c->dsys[sys].nb_tab_off + t] = nh;
    HIPCHK(c, memcpy_st(c, c->nb_tabs.p + c->dsys[sys].nb_tab_off + t, &nh, sizeof nh, hipMemcpyHostToDevice));
    return 0;
}

int ibx_set_thermostat(ibx_ctx *c, int ensemble, double temp_in, double fac, double dt) {
    if (!c || (ensemble != 0 && ensemble != 1)) return fail(c, IBX_E_INVALID, "invalid thermostat");
    HIPCHK(c, hipStreamSynchronize(c->st));
    // the device copy carries the tuned skins: start from it
    HIPCHK(c, memcpy_st(c, c->dsys.data(), c->sys.p, c->dsys.size() * sizeof(DevSys), hipMemcpyDeviceToHost));
    for (auto &D : c->dsys) {
        D.ensemble = ensemble;
        D.temp_in = temp_in;
        D.fac = fac;
        D.dt = dt;
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, memcpy_st(c, c->sys.p, c->dsys.data(), c->dsys.size() * sizeof(DevSys), hipMemcpyHostToDevice));
    drop_graph(c);
    return 0;
}

int ibx_get_pairs(ibx_ctx *c, int sys, int group, int *out, int64_t cap, int64_t *count) {
    if (!c || sys < 0 || sys >= c->nsys || group < 0 || group > 1 || !count)
        return fail(c, IBX_E_INVALID, "invalid arguments");
    *count = 0;
    GroupDev &G = c->gr[group];
    if (!G.present) return 0;
    const DevGroup &dg = c->dsys[sys].g[group];
    if (dg.nmem == 0) return 0;
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::vector<int2> nneb(dg.nslot_pad);
    std::vector<int> perm(dg.nslot_pad);
    HIPCHK(c, memcpy_st(c, nneb.data(), G.nneb.p + dg.slot_off, nneb.size() * sizeof(int2), hipMemcpyDeviceToHost));
    HIPCHK(c, memcpy_st(c, perm.data(), G.perm.p + dg.slot_off, perm.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (!out) {   // rows are full (each pair in both members' rows): the count is half the row lengths
        int64_t n = 0;
        for (int i = 0; i < dg.nslot_pad; ++i) {
            if (perm[i] < 0) continue;
            const int n_in = std::min(nneb[i].x, G.cap);
            n += n_in + std::min(nneb[i].y, G.cap - n_in);
        }
        *count = n / 2;
        return 0;
    }
    std::vector<int> list((size_t)dg.nslot_pad * G.cap);
    HIPCHK(c, memcpy_st(c, list.data(), G.list.p + (size_t)dg.slot_off * G.cap, list.size() * sizeof(int),
                        hipMemcpyDeviceToHost));
    const int o = c->item_off[sys];
    std::vector<std::pair<int, int>> pr;
    for (int i = 0; i < dg.nslot_pad; ++i) {
        if (perm[i] < 0) continue;
        const int I = perm[i] - o;
        const int n_in = std::min(nneb[i].x, G.cap), n_out = std::min(nneb[i].y, G.cap - n_in);
        for (int k = 0; k < n_in + n_out; ++k) {
            const int j = list[list_pos(i, k < n_in ? k : G.cap - 1 - (k - n_in), G.cap)] - dg.slot_off;
            const int J = perm[j] - o;
            if (I < J) pr.emplace_back(I, J);
        }
    }
    std::sort(pr.begin(), pr.end());
    *count = (int64_t)pr.size();
    if (out && cap >= *count)
        for (size_t q = 0; q < pr.size(); ++q) { out[2 * q] = pr[q].first; out[2 * q + 1] = pr[q].second; }
    return 0;
}

int ibx_rdf_enable(ibx_ctx *c, int nbin, double dr, int every) {
    if (!c || nbin < 0 || nbin > 8192 || (nbin > 0 && !(dr > 0))) return fail(c, IBX_E_INVALID, "invalid RDF settings");
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->rdf_nbin = nbin;
    c->rdf_dr = dr;
    c->rdf_every = every;
    if (nbin) {
        HIPCHK(c, c->rdf.alloc((size_t)c->nsys * nbin));
        HIPCHK(c, memset_st(c, c->rdf.p, 0, c->rdf.bytes()));
        HIPCHK(c, c->rdf_samples.alloc(c->nsys));
        HIPCHK(c, memset_st(c, c->rdf_samples.p, 0, c->rdf_samples.bytes()));
        HIPCHK(c, c->rdf_ivol.alloc(c->nsys));
        HIPCHK(c, memset_st(c, c->rdf_ivol.p, 0, c->rdf_ivol.bytes()));
    }
    drop_graph(c);
    return 0;
}

int ibx_rdf_get(ibx_ctx *c, int sys, double *g, int reset) {
    if (!c || sys < 0 || sys >= c->nsys || !c->rdf_nbin) return fail(c, IBX_E_INVALID, "RDF not enabled");
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::vector<unsigned long long> h(c->rdf_nbin);
    HIPCHK(c, memcpy_st(c, h.data(), c->rdf.p + (size_t)sys * c->rdf_nbin, h.size() * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost));
    double ivol = 0.0;
    HIPCHK(c, memcpy_st(c, &ivol, c->rdf_ivol.p + sys, sizeof(double), hipMemcpyDeviceToHost));
    const HostSystem &hs = c->hs[sys];
    const double n = (double)hs.members[0].size();
    // samples at volumes V_i expect n (n / V_i) shell pairs each; no sample yet: an empty histogram
    if (!(ivol > 0.0)) ivol = 1.0 / (hs.box[0] * hs.box[1] * hs.box[2]);
    for (int b = 0; b < c->rdf_nbin; ++b) {
        const double r0 = b * c->rdf_dr, r1 = (b + 1) * c->rdf_dr;
        const double shell = 4.0 / 3.0 * M_PI * (r1 * r1 * r1 - r0 * r0 * r0);
        g[b] = (double)h[b] / (n * n * ivol * shell);
    }
    if (reset) {
        HIPCHK(c, memset_st(c, c->rdf.p + (size_t)sys * c->rdf_nbin, 0, c->rdf_nbin * sizeof(unsigned long long)));
        HIPCHK(c, memset_st(c, c->rdf_samples.p + sys, 0, sizeof(int)));
        HIPCHK(c, memset_st(c, c->rdf_ivol.p + sys, 0, sizeof(double)));
    }
    return 0;
}

int ibx_get_timing(ibx_ctx *c, double ms[8]) {
    if (!c) return fail(nullptr, IBX_E_INVALID, "null context");
    for (int k = 0; k < 8; ++k) ms[k] = c->tm.ms[k];
    return 0;
}

int ibx_get_info(ibx_ctx *c, int sys, int64_t info[8]) {
    if (!c || sys < 0 || sys >= c->nsys) return fail(c, IBX_E_INVALID, "invalid system index");
    const HostSystem &h = c->hs[sys];
    const size_t bytes = device_bytes(c);
    info[0] = h.natoms;
    info[1] = h.nvs;
    info[2] = c->gr[0].cap;
    info[3] = c->gr[1].cap;
    info[4] = c->dsys[sys].g[0].nslot_pad;
    info[5] = (int64_t)bytes;
    info[6] = c->nsys;
    info[7] = c->n_items;
    return 0;
}

// Enables per-kernel-class device timing of ibx_step (off by default).
int ibx_set_timing(ibx_ctx *c, int on) {
    if (!c) return fail(nullptr, IBX_E_INVALID, "null context");
    c->timing = on != 0;
    for (double &x : c->tm.ms) x = 0.0;
    return 0;
}

int ibx_load_dir(const char *dir, ibx_ctx **out) {
    DescStorage st;
    std::string err;
    int rc = read_input_dir(dir, st, err);
    if (rc) return fail(nullptr, rc, "%s", err.c_str());
    const ibx_system_desc *d = &st.d;
    return ibx_create(1, &d, out);
}

}  // extern "C"
