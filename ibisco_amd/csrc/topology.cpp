// Host-side set-up: tables, bonded-term enumeration, exclusions, groups.
//
// Follows the reference's set-up semantics (cited per function) with O(N)
// data structures; nothing here runs per step.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "ibx_internal.h"

namespace ibx {

namespace {

const double kPi = 3.14159265358;   // module_var.F90:16

// Table load (RDINTERACT.f90:205-220 for bonds: rows as read; 450-537 for the
// other kinds: a table whose first abscissa is not 0 gets a row at 0 and a row
// at one bin, both carrying the first energy).
void load_table(Table &t, int nrow, const double *x, const double *v, bool shift, bool oop) {
    t.x.assign(nrow + 4, 0.0);
    t.V.assign(nrow + 4, 0.0);
    t.F.assign(nrow + 4, 0.0);
    double bin = x[1] - x[0];
    if (oop && x[1] < x[0]) bin = x[0] - x[1];
    int k;
    if (shift && x[0] != 0.0) {
        t.x[0] = 0.0; t.V[0] = v[0];
        t.x[1] = bin; t.V[1] = v[0];
        t.x[2] = x[1]; t.V[2] = v[1];
        for (k = 3; k - 1 < nrow; ++k) { t.x[k] = x[k - 1]; t.V[k] = v[k - 1]; }
    } else {
        for (k = 0; k < nrow; ++k) { t.x[k] = x[k]; t.V[k] = v[k]; }
    }
    t.ndat = k - 1;
    t.bin = bin;
    t.x0 = t.x[0];
}

// FTABLE.f90:9-91.  kind: 0 bond (F/r), 1 angle, 2 torsion, 3 nonbonded (F/r), 4 OOP.
void differentiate(Table &t, int kind) {
    const double d2r = kPi / 180.0;
    int n = t.ndat;
    if (kind == 3) {
        for (int j = 1; j <= n; ++j) t.F[j] = -(t.V[j + 1] - t.V[j - 1]) / t.bin / t.x[j] / 2.0;
        t.F[0] = t.F[1];
        return;
    }
    for (int j = 1; j <= n - 1; ++j) {
        double dv = -(t.V[j + 1] - t.V[j - 1]) / t.bin;
        t.F[j] = (kind == 0 ? dv / t.x[j] : dv / d2r) / 2.0;
    }
    t.F[0] = t.F[1];
    // bonds/angles repeat the previous row at the end (FTABLE.f90:31,38);
    // torsions and OOP wrap to row 0 (FTABLE.f90:49,81)
    t.F[n] = (kind == 2 || kind == 4) ? t.F[0] : t.F[n - 1];
}

int type_class(const std::vector<int> &tlabel, const int *types, int k) {
    int h = 0;
    for (int i = 0; i < k; ++i) h += tlabel[types[i]];
    if (h == k) return 1;
    if (h == 2 * k) return 2;
    return 3;
}

struct TypeMap {
    int nt = 0, k = 0;
    std::vector<int> m;
    int at(const int *t) const {
        size_t idx = 0;
        for (int i = 0; i < k; ++i) idx = idx * nt + t[i];
        return m[idx];
    }
};

int setup_kind(HostSystem &hs, const ibx_tables &src, int k, int kind, std::vector<Table> &dst,
               TypeMap &map, std::vector<int> *labels, bool symmetric, std::string &err) {
    map.nt = hs.ntype;
    map.k = k;
    size_t sz = 1;
    for (int i = 0; i < k; ++i) sz *= (size_t)hs.ntype;
    map.m.assign(sz, -1);
    dst.resize(src.ntab);
    if (labels) labels->assign(src.ntab, 1);
    for (int t = 0; t < src.ntab; ++t) {
        const int *ty = src.types + t * k;
        size_t fwd = 0, bwd = 0;
        for (int i = 0; i < k; ++i) {
            if (ty[i] < 0 || ty[i] >= hs.ntype) { err = "table type out of range"; return IBX_E_INVALID; }
            fwd = fwd * hs.ntype + ty[i];
        }
        for (int i = k - 1; i >= 0; --i) bwd = bwd * hs.ntype + ty[i];
        if (map.m[fwd] != -1) {
            err = "duplicate entry in interaction file";   // RDINTERACT.f90:186-194
            return IBX_E_INVALID;
        }
        map.m[fwd] = t;
        if (symmetric) map.m[bwd] = t;
        if (src.nrow[t] < 3) { err = "table with fewer than 3 rows"; return IBX_E_INVALID; }
        load_table(dst[t], src.nrow[t], src.x + src.off[t], src.v + src.off[t], kind != 0, kind == 4);
        differentiate(dst[t], kind);
        if (labels) (*labels)[t] = type_class(hs.tlabel, ty, k);
    }
    return 0;
}

}  // namespace

int build_host_system(const ibx_system_desc *d, HostSystem &hs, std::string &err) {
    if (d->natoms < 2 || d->ntype < 1 || d->nvs < 0) { err = "invalid sizes"; return IBX_E_INVALID; }
    hs.natoms = d->natoms;
    hs.nvs = d->nvs;
    hs.nitems = d->natoms + d->nvs;
    hs.ntype = d->ntype;
    hs.hybrid = d->nvs > 0;
    const int n = hs.nitems, na = hs.natoms, nt = hs.ntype;
    hs.mass.assign(d->mass, d->mass + nt);
    hs.invmass.resize(nt);
    hs.tlabel.assign(d->label, d->label + nt);
    for (int t = 0; t < nt; ++t) {
        if (!(hs.mass[t] > 0)) { err = "non-positive mass"; return IBX_E_INVALID; }
        hs.invmass[t] = 1.0 / hs.mass[t];
        if (hs.tlabel[t] != 1 && hs.tlabel[t] != 2) { err = "the label must be equal to A or B"; return IBX_E_INVALID; }
    }
    hs.itype.assign(n, 0);
    hs.type_label.assign(n, 1);
    hs.nbonds.assign(d->nbonds, d->nbonds + na);
    hs.jbond.assign(d->jbond, d->jbond + 4 * (size_t)na);
    hs.s.assign(3 * (size_t)n, 0.0);
    std::copy(d->s, d->s + 3 * (size_t)na, hs.s.begin());
    hs.v.assign(d->v, d->v + 3 * (size_t)na);
    for (int i = 0; i < na; ++i) {
        if (d->itype[i] < 0 || d->itype[i] >= nt) { err = "atom type out of range"; return IBX_E_INVALID; }
        hs.itype[i] = d->itype[i];
        hs.type_label[i] = hs.hybrid ? hs.tlabel[d->itype[i]] : 1;   // RDCOOR.f90:84,110-123
        if (hs.nbonds[i] < 0 || hs.nbonds[i] > 4) { err = "more than 4 bonds on an atom"; return IBX_E_INVALID; }
        for (int b = 0; b < hs.nbonds[i]; ++b)
            if (hs.jbond[4 * i + b] < 0 || hs.jbond[4 * i + b] >= na) { err = "bond partner out of range"; return IBX_E_INVALID; }
    }
    for (int k = 0; k < 3; ++k) {
        hs.box[k] = d->box[k];
        if (!(hs.box[k] > 0)) { err = "box length must be positive"; return IBX_E_INVALID; }
    }
    TypeMap mb, ma, mt, mn, mo;
    int rc;
    if ((rc = setup_kind(hs, d->bond, 2, 0, hs.bt, mb, &hs.bond_lab, true, err)) ||
        (rc = setup_kind(hs, d->angle, 3, 1, hs.at, ma, &hs.angle_lab, true, err)) ||
        (rc = setup_kind(hs, d->tors, 4, 2, hs.tt, mt, &hs.tors_lab, true, err)) ||
        (rc = setup_kind(hs, d->nb, 2, 3, hs.nt, mn, nullptr, true, err)) ||
        (rc = setup_kind(hs, d->oop, 4, 4, hs.ot, mo, &hs.oop_lab, false, err)))
        return rc;
    hs.inbondt = mn.m;

    // virtual sites (RDVIRTUAL.f90:40-140)
    hs.vs_of_atom.assign(na, -1);
    if (hs.hybrid) {
        hs.vs_type.assign(d->vs_type, d->vs_type + hs.nvs);
        hs.vs_center.assign(d->vs_center, d->vs_center + hs.nvs);
        hs.vs_off.assign(d->vs_off, d->vs_off + hs.nvs + 1);
        hs.vs_atoms.assign(d->vs_atoms, d->vs_atoms + hs.vs_off[hs.nvs]);
        hs.vs_invmass.resize(hs.nvs);
        hs.vs_coeff.resize(hs.vs_atoms.size());
        for (int i = 0; i < hs.nvs; ++i) {
            int pos = na + i;
            hs.itype[pos] = hs.vs_type[i];
            hs.type_label[pos] = 3;
            double m = 0.0;
            for (int p = hs.vs_off[i]; p < hs.vs_off[i + 1]; ++p) {
                int a = hs.vs_atoms[p];
                if (a < 0 || a >= na) { err = "virtual-site member out of range"; return IBX_E_INVALID; }
                m += hs.mass[hs.itype[a]];
                // DISTRIBUTE_VSFORCE.f90:21-38 assumes each atom belongs to one VS
                if (hs.vs_of_atom[a] != -1) { err = "atom belongs to more than one virtual site"; return IBX_E_INVALID; }
                hs.vs_of_atom[a] = i;
            }
            hs.vs_invmass[i] = 1.0 / m;
            for (int p = hs.vs_off[i]; p < hs.vs_off[i + 1]; ++p) hs.vs_coeff[p] = hs.mass[hs.itype[hs.vs_atoms[p]]] / m;
        }
    }

    // ATOM / BEAD lists (MAKE_LISTS.f90:13-60)
    hs.members[0].clear();
    hs.members[1].clear();
    for (int i = 0; i < na; ++i) (hs.type_label[i] == 1 ? hs.members[0] : hs.members[1]).push_back(i);
    hs.numatoms = (int)hs.members[0].size();
    hs.numbeads = (int)hs.members[1].size();
    for (int i = 0; i < hs.nvs; ++i) hs.members[1].push_back(na + i);
    hs.ngroup = hs.hybrid ? 2 : 1;
    hs.rcut[0] = d->rcut_atom; hs.rlist[0] = d->rlist_atom;
    hs.rcut[1] = d->rcut_bead; hs.rlist[1] = d->rlist_bead;
    for (int g = 0; g < hs.ngroup; ++g)
        if (!(hs.rlist[g] >= hs.rcut[g]) || !(hs.rcut[g] > 0)) {
            err = "neighbour list cutoff must be larger than the cutoff";   // RDCONTROL.f90:375
            return IBX_E_INVALID;
        }
    for (int g = 0; g < hs.ngroup; ++g) {
        std::vector<char> present(nt, 0);
        for (int i : hs.members[g]) present[hs.itype[i]] = 1;
        for (int a = 0; a < nt; ++a)
            for (int b = 0; b < nt; ++b)
                if (present[a] && present[b] && hs.inbondt[a * nt + b] < 0) {
                    err = "missing nonbonded table for a type pair in use";
                    return IBX_E_INVALID;
                }
    }

    // Terms (SETLIS.f90:26-419): angle i-j-k owned by i with k>i, torsion
    // i-j-k-l owned by i with l>i; parameterless terms only exclude.
    struct Own { int o, a, b, c; };
    std::vector<Own> ang, noang, tor, ftor, oop;
    if (!hs.bt.empty()) {
        for (int I = 0; I < na; ++I) {
            const int IT = hs.itype[I];
            const int *jb = &hs.jbond[4 * (size_t)I];
            if (!hs.ot.empty() && hs.nbonds[I] == 3) {
                static const int perm[6][3] = {{0, 1, 2}, {2, 0, 1}, {1, 0, 2}, {0, 2, 1}, {1, 2, 0}, {2, 1, 0}};
                bool found = false;
                for (int p = 0; p < 6 && !found; ++p) {
                    int Z = jb[perm[p][0]], P = jb[perm[p][1]], O = jb[perm[p][2]];
                    int ty[4] = {IT, hs.itype[Z], hs.itype[P], hs.itype[O]};
                    if (mo.at(ty) != -1) { oop.push_back({I, Z, P, O}); found = true; }
                }
            }
            for (int j1 = 0; j1 < hs.nbonds[I]; ++j1) {
                const int J = jb[j1], JT = hs.itype[J];
                int tb[2] = {IT, JT};
                if (mb.at(tb) == -1) {
                    char buf[96];
                    snprintf(buf, sizeof buf, "no bond parameters for atoms %d and %d", I + 1, J + 1);
                    err = buf;                      // SETLIS.f90:105-124
                    return IBX_E_INVALID;
                }
                if (hs.at.empty()) continue;
                for (int k1 = 0; k1 < hs.nbonds[J]; ++k1) {
                    const int K = hs.jbond[4 * (size_t)J + k1], KT = hs.itype[K];
                    if (K == I) continue;
                    if (K > I) {
                        int ta[3] = {IT, JT, KT};
                        (ma.at(ta) == -1 ? noang : ang).push_back({I, J, K, -1});
                    }
                    if (hs.tt.empty()) continue;
                    for (int l1 = 0; l1 < hs.nbonds[K]; ++l1) {
                        const int L = hs.jbond[4 * (size_t)K + l1], LT = hs.itype[L];
                        if (L == J || L <= I) continue;
                        int t4[4] = {IT, JT, KT, LT};
                        if (mt.at(t4) != -1) tor.push_back({I, J, K, L});
                        else if (!hs.ot.empty() && mo.at(t4) != -1) oop.push_back({I, J, K, L});
                        else ftor.push_back({I, J, K, L});
                    }
                }
            }
        }
    }
    // virtual angles join the owner's list after SETLIS (RDVIRTANGLES.f90:36-50)
    for (int q = 0; q < d->nvirtangles; ++q) {
        const int *va = d->virtangles + 3 * q;
        for (int k = 0; k < 3; ++k)
            if (va[k] < 0 || va[k] >= n) { err = "virtangles index out of range"; return IBX_E_INVALID; }
        ang.push_back({va[0], va[1], va[2], -1});
    }
    auto by_owner = [](std::vector<Own> &v) {
        std::stable_sort(v.begin(), v.end(), [](const Own &x, const Own &y) { return x.o < y.o; });
    };
    by_owner(ang); by_owner(oop);

    // Bonds owned by the lower index (BONDS.inc:11), in owner order.
    hs.bonds.clear(); hs.bond_tab.clear();
    for (int I = 0; I < na; ++I)
        for (int b = 0; b < hs.nbonds[I]; ++b) {
            int J = hs.jbond[4 * (size_t)I + b];
            if (J > I) {
                int tb[2] = {hs.itype[I], hs.itype[J]};
                hs.bonds.push_back({I, J, -1, -1});
                hs.bond_tab.push_back(mb.at(tb));
            }
        }
    hs.angles.clear(); hs.angle_tab.clear();
    for (auto &t : ang) {
        int ta[3] = {hs.itype[t.o], hs.itype[t.a], hs.itype[t.b]};
        int tab = ma.at(ta);
        if (tab < 0) { err = "virtual angle with no angle table"; return IBX_E_INVALID; }
        hs.angles.push_back({t.o, t.a, t.b, -1});
        hs.angle_tab.push_back(tab);
    }
    hs.tors.clear(); hs.tors_tab.clear();
    for (auto &t : tor) {
        int t4[4] = {hs.itype[t.o], hs.itype[t.a], hs.itype[t.b], hs.itype[t.c]};
        hs.tors.push_back({t.o, t.a, t.b, t.c});
        hs.tors_tab.push_back(mt.at(t4));
    }
    hs.oops.clear(); hs.oop_tab.clear();
    for (auto &t : oop) {
        int t4[4] = {hs.itype[t.o], hs.itype[t.a], hs.itype[t.b], hs.itype[t.c]};
        hs.oops.push_back({t.o, t.a, t.b, t.c});
        hs.oop_tab.push_back(mo.at(t4));
    }

    // Exclusions (BUILD_CONNECTIVITY.f90:14-116 then RDVIRTBONDS.f90:42-64):
    // partners over atoms from bonds, angles (with and without parameters) and,
    // for groups with exclusion 5, torsions (with and without parameters) and
    // OOP terms; each list ascending, then the virtbonds pairs appended.
    std::vector<std::vector<int>> base(n), p14(n);
    auto link = [&](std::vector<std::vector<int>> &v, int a, int b) {
        if (a < na && b < na) { v[a].push_back(b); v[b].push_back(a); }
    };
    for (int I = 0; I < na; ++I)
        for (int b = 0; b < hs.nbonds[I]; ++b) link(base, I, hs.jbond[4 * (size_t)I + b]);
    for (auto &t : ang) if (t.o < na) { link(base, t.o, t.a); link(base, t.o, t.b); }
    for (auto &t : noang) { link(base, t.o, t.a); link(base, t.o, t.b); }
    for (auto *lst : {&tor, &ftor, &oop})
        for (auto &t : *lst) { link(p14, t.o, t.a); link(p14, t.o, t.b); link(p14, t.o, t.c); }
    const int excl_of_group[2] = {d->nonbexc_atom, d->nonbexc_bead};
    for (int g = 0; g < 2; ++g)
        if (excl_of_group[g] != 4 && excl_of_group[g] != 5 && (g == 0 || hs.hybrid)) {
            err = "non_bonded must be 4 or 5";   // RDCONTROL.f90:435
            return IBX_E_INVALID;
        }
    std::vector<std::vector<int>> conn(n);
    for (int I = 0; I < na; ++I) {
        int g = hs.type_label[I] == 2 ? 1 : 0;
        std::vector<int> c = base[I];
        if (excl_of_group[g] == 5) c.insert(c.end(), p14[I].begin(), p14[I].end());
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
        c.erase(std::remove(c.begin(), c.end(), I), c.end());
        conn[I] = c;
    }
    for (int q = 0; q < d->nvirtbonds; ++q) {
        int I = d->virtbonds[2 * q], J = d->virtbonds[2 * q + 1];
        if (I < 0 || I >= n || J < 0 || J >= n) { err = "virtbonds index out of range"; return IBX_E_INVALID; }
        conn[I].push_back(J);
        conn[J].push_back(I);
    }
    hs.excl_off.assign(n + 1, 0);
    for (int I = 0; I < n; ++I) {
        if (conn[I].size() > 20) { err = "more than MAXCONNECTIONS=20 exclusions"; return IBX_E_INVALID; }
        hs.excl_off[I + 1] = hs.excl_off[I] + (int)conn[I].size();
    }
    hs.excl.clear();
    for (int I = 0; I < n; ++I) hs.excl.insert(hs.excl.end(), conn[I].begin(), conn[I].end());

    hs.ensemble = d->ensemble;
    hs.dt = d->dt; hs.temp_in = d->temp_in; hs.fac = d->fac;
    hs.nupdate = d->nupdate; hs.halt_drift = d->halt_drift; hs.nmts = d->nmts; hs.mts_quirk = d->mts_quirk;
    if (hs.nupdate <= 0 || hs.halt_drift == 0) { err = "update_neighbour_list and halt_drift must be positive"; return IBX_E_INVALID; }
    if (hs.ensemble < 0 || hs.ensemble > 2) { err = "ensemble must be NVE, NVT or NPT"; return IBX_E_INVALID; }
    hs.pressure = d->pressure;
    hs.pfac = d->pfac;
    if (hs.ensemble == 2 && !(hs.pfac >= 0.0)) { err = "NPT needs a non-negative BETA*DT*LIMAVP/TAUP"; return IBX_E_INVALID; }

    // initial temperatures (main.F90:173-195) drive the first Berendsen factor
    hs.mktemp = 2.0 / (3.0 * (na - 1));
    hs.mktemp_atom = 2.0 / (3.0 * (hs.numatoms - 1));
    hs.mktemp_bead = 2.0 / (3.0 * (hs.numbeads - 1));
    double ek[2] = {0, 0};
    for (int g = 0; g < 2; ++g)
        for (int i : hs.members[g]) {
            if (i >= na) continue;
            const double *vi = &hs.v[3 * (size_t)i];
            ek[g] += hs.mass[hs.itype[i]] * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
        }
    hs.temp_atom0 = 0.5 * ek[0] * hs.mktemp_atom;
    hs.temp_bead0 = 0.5 * ek[1] * hs.mktemp_bead;
    return 0;
}

}  // namespace ibx
