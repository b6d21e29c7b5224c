// Readers of the reference's input files for ibx_load_dir():
// control (RDCONTROL.f90), interaction and its table files (RDINTERACT.f90),
// coordinate (RDCOOR.f90), virtual / virtbonds / virtangles (RDVIRTUAL.f90,
// RDVIRTBONDS.f90, RDVIRTANGLES.f90).  Units are converted to the reduced set
// of UNIT.f90 in FP64.
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "ibx_internal.h"

namespace ibx {

namespace {

const double kKB = 1.380658e-23, kNA = 6.0221367e23, kMMOLY = 0.012;   // module_var.F90:14-17

struct Units {
    double massscale = kMMOLY / kNA, rscale = 1.0e-9, escale = kKB * 300.0;
    double timescale = rscale * std::sqrt(massscale / escale);
    double vscale = rscale / timescale;
    double tempscale = escale / kKB;
    double pscale = escale / (rscale * rscale * rscale) / 1000.0;   // reduced pressure -> kPa
};

std::vector<std::string> tokens(const std::string &line) {
    std::string l = line.substr(0, line.find('#'));
    std::istringstream is(l);
    std::vector<std::string> t;
    std::string w;
    while (is >> w) t.push_back(w);
    return t;
}

bool to_num(std::string s, double &x) {
    for (char &ch : s)
        if (ch == 'D' || ch == 'd') ch = 'E';
    char *end = nullptr;
    x = std::strtod(s.c_str(), &end);
    return end && *end == '\0' && end != s.c_str();
}

bool read_lines(const std::string &path, std::vector<std::string> &out) {
    std::ifstream f(path);
    if (!f) return false;
    std::string l;
    while (std::getline(f, l)) out.push_back(l);
    return true;
}

// Two-column table file; reading stops at the first line that is not two numbers.
bool read_table(const std::string &path, std::vector<double> &x, std::vector<double> &v) {
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) return false;
    for (auto &l : lines) {
        auto t = tokens(l);
        double a, b;
        if (t.size() < 2 || !to_num(t[0], a) || !to_num(t[1], b)) break;
        x.push_back(a);
        v.push_back(b);
    }
    return x.size() >= 3;
}

}  // namespace

int read_input_dir(const char *dir, DescStorage &st, std::string &err) {
    const std::string d = std::string(dir) + "/";
    const Units U;
    const double econv = 1000.0 / kNA / U.escale;

    // control: the first line naming a keyword wins, except mts where the last does
    std::vector<std::string> cl;
    if (!read_lines(d + "control", cl)) { err = "File control does not exist"; return IBX_E_IO; }
    std::map<std::string, std::string> kv;
    for (auto &l : cl) {
        auto t = tokens(l);
        if (t.size() < 2) continue;
        if (t[0] == "mts" || !kv.count(t[0])) kv[t[0]] = t[1];
    }
    auto num = [&](const char *k, double def, bool required) -> double {
        auto it = kv.find(k);
        double x;
        if (it == kv.end() || !to_num(it->second, x)) {
            if (required) err += std::string("missing control keyword ") + k + "; ";
            return def;
        }
        return x;
    };
    ibx_system_desc &D = st.d;
    std::string ens = kv.count("ensemble") ? kv["ensemble"] : "";
    for (auto &ch : ens) ch = (char)toupper(ch);
    if (ens == "NVE") D.ensemble = 0;
    else if (ens == "NVT") D.ensemble = 1;
    else if (ens == "NPT") D.ensemble = 2;
    else { err = "You did not input a valid Ensemble in control (NVE/NVT/NPT on the device path)"; return IBX_E_INVALID; }
    const double temp_k = num("temperature", 300, true);
    const int natoms = (int)num("atoms", 0, true);
    st.nstep = (int)num("num_of_time_steps", 0, true);
    const double dt_ps = num("time_step", 0, true);
    const int nvs = (int)num("virtual_sites", 0, true);
    D.rcut_atom = num("cutoff", 0, true);
    D.rlist_atom = num("neighbour_list_cutoff", 0, true);
    D.rcut_bead = num("bead_cutoff", D.rcut_atom, nvs > 0);
    D.rlist_bead = num("bead_neighbour_list_cutoff", D.rlist_atom, nvs > 0);
    D.nonbexc_atom = (int)num("non_bonded", 4, true);
    D.nonbexc_bead = (int)num("non_bonded_bead", 4, nvs > 0);
    D.nupdate = (int)num("update_neighbour_list", 0, true);
    D.halt_drift = (int)num("halt_drift", 0, true);
    st.ntrj = (int)num("trajectory", 0, true);
    const double taut = num("temperature_coupling_time", 0.1, true);
    // RDCONTROL.f90:99-108, 307-345: the barostat's keywords are required for NPT
    const double pressure_kpa = num("pressure", 101.3, D.ensemble == 2);
    const double taup = num("pressure_coupling_time", 5.0, D.ensemble == 2);
    const double beta = num("isothermal_compressibility", 1.0e-6, D.ensemble == 2);
    D.nmts = (int)num("mts", 0, true);
    if (!err.empty()) return IBX_E_INVALID;
    D.dt = dt_ps * 1.0e-12 / U.timescale;
    D.temp_in = temp_k / U.tempscale;
    D.fac = D.dt / (taut * 1.0e-12 / U.timescale);
    D.pressure = pressure_kpa / U.pscale;
    D.pfac = beta * U.pscale * D.dt * 10.0 / (taup * 1.0e-12 / U.timescale);   // ALLOCATEVAR.f90:17, LIMAVP = 10
    // the reference's MTS_LOOP takes I = MOD(STEP, 3+NMTS) = 0 on the last step of each cycle
    // (MTS_LOOP.f90:100, :176): an input directory runs the same cycle
    D.mts_quirk = 1;

    // interaction
    std::vector<std::string> il;
    if (!read_lines(d + "interaction", il)) { err = "File interaction does not exist"; return IBX_E_IO; }
    size_t p = 0;
    auto next = [&]() -> std::vector<std::string> {
        while (p < il.size()) {
            auto t = tokens(il[p++]);
            if (!t.empty()) return t;
        }
        return {};
    };
    auto t = next();
    if (t.size() < 2 || t[0] != "atom_types") { err = "interaction: atom_types expected"; return IBX_E_INVALID; }
    const int ntype = std::atoi(t[1].c_str());
    st.mass.assign(ntype, 0.0);
    st.label.assign(ntype, 1);
    for (int k = 0; k < ntype; ++k) {
        t = next();
        double m;
        if (t.size() < 4 || !to_num(t[2], m)) { err = "Number of atom_types different from Number of Atom"; return IBX_E_INVALID; }
        int ty = std::atoi(t[0].c_str()) - 1;
        if (ty < 0 || ty >= ntype) { err = "atom type index out of range"; return IBX_E_INVALID; }
        st.mass[ty] = m / kNA / U.massscale;
        char lab = (char)toupper(t[3][0]);
        if (lab != 'A' && lab != 'B') { err = "the label for hybrid model must be equal to A or B"; return IBX_E_INVALID; }
        st.label[ty] = lab == 'A' ? 1 : 2;
    }
    const char *sect[5] = {"bond_types", "angle_types", "torsion_types", "non_bonded_interaction_types", "Out_of_Planes"};
    const int arity[5] = {2, 3, 4, 2, 4};
    for (int k = 0; k < 5; ++k) {
        t = next();
        if (t.size() < 2 || t[0] != sect[k]) { err = std::string("interaction: section ") + sect[k] + " expected"; return IBX_E_INVALID; }
        const int n = std::atoi(t[1].c_str());
        for (int q = 0; q < n; ++q) {
            t = next();
            if ((int)t.size() < arity[k] + 1) { err = std::string("interaction: short line in ") + sect[k]; return IBX_E_INVALID; }
            for (int a = 0; a < arity[k]; ++a) st.ttypes[k].push_back(std::atoi(t[a].c_str()) - 1);
            std::vector<double> x, v;
            if (!read_table(d + t[arity[k]], x, v)) { err = "File " + t[arity[k]] + " does not exist or is too short"; return IBX_E_IO; }
            st.toff[k].push_back((int)st.tx[k].size());
            st.tnrow[k].push_back((int)x.size());
            for (size_t r = 0; r < x.size(); ++r) {
                st.tx[k].push_back(x[r]);
                st.tv[k].push_back(v[r] * econv);
            }
        }
    }

    // coordinate
    std::vector<std::string> co;
    if (!read_lines(d + "coordinate", co)) { err = "File coordinate does not exist"; return IBX_E_IO; }
    if (co.size() < 10) { err = "Fatal error in RDCOOR, check coordinate file"; return IBX_E_INVALID; }
    {
        auto bt = tokens(co[3]);
        if (bt.size() < 3 || !to_num(bt[0], D.box[0]) || !to_num(bt[1], D.box[1]) || !to_num(bt[2], D.box[2])) {
            err = "Fatal error in RDCOOR, check coordinate file (box)";
            return IBX_E_INVALID;
        }
    }
    std::vector<std::string> tk;
    for (size_t l = 9; l < co.size(); ++l) {
        auto tt = tokens(co[l]);
        tk.insert(tk.end(), tt.begin(), tt.end());
    }
    size_t q = 0;
    if (tk.size() < 2) { err = "Fatal error in RDCOOR, check coordinate file"; return IBX_E_INVALID; }
    const int nmol = std::atoi(tk[1].c_str());
    q = 2;
    st.itype.clear();
    for (int m = 0; m < nmol; ++m) {
        if (q + 4 > tk.size()) { err = "Some molecules name missing. Check your coordinate file"; return IBX_E_INVALID; }
        const int na = std::atoi(tk[q].c_str());
        q += 4;
        for (int a = 0; a < na; ++a) {
            if (q + 6 > tk.size()) { err = "No. Atoms in coordinate does not equal to NATOMS"; return IBX_E_INVALID; }
            const int ty = std::atoi(tk[q + 1].c_str()) - 1, nb = std::atoi(tk[q + 2].c_str());
            double xyz[3];
            for (int k = 0; k < 3; ++k)
                if (!to_num(tk[q + 3 + k], xyz[k])) { err = "Fatal error in RDCOOR, check coordinate file"; return IBX_E_INVALID; }
            q += 6;
            double tmp;
            while (q < tk.size() && !to_num(tk[q], tmp)) ++q;   // trailing labels on the position line
            double vel[3];
            for (int k = 0; k < 3; ++k)
                if (q + k >= tk.size() || !to_num(tk[q + k], vel[k])) { err = "Fatal error in RDCOOR, check coordinate file"; return IBX_E_INVALID; }
            q += 3;
            if (nb < 0 || nb > 4 || q + nb > tk.size()) { err = "Fatal error in RDCOOR, check bonds"; return IBX_E_INVALID; }
            st.itype.push_back(ty);
            st.nbonds.push_back(nb);
            for (int b = 0; b < 4; ++b) st.jbond.push_back(b < nb ? std::atoi(tk[q + b].c_str()) - 1 : -1);
            q += nb;
            for (int k = 0; k < 3; ++k) {
                st.s.push_back(xyz[k]);
                st.v.push_back(vel[k] * 1.0e3 / U.vscale);
            }
        }
    }
    if ((int)st.itype.size() != natoms) { err = "No. Atoms in coordinate does not equal to NATOMS"; return IBX_E_INVALID; }

    // hybrid inputs
    if (nvs > 0) {
        std::vector<std::string> vl;
        if (!read_lines(d + "virtual", vl)) { err = "File virtual does not exist"; return IBX_E_IO; }
        std::vector<std::string> vt;
        for (size_t l = 5; l < vl.size(); ++l) {
            auto tt = tokens(vl[l]);
            vt.insert(vt.end(), tt.begin(), tt.end());
        }
        size_t r = 0;
        st.vs_off.push_back(0);
        for (int v = 0; v < nvs; ++v) {
            if (r + 4 > vt.size()) { err = "virtual: too few records"; return IBX_E_INVALID; }
            const int ty = std::atoi(vt[r + 1].c_str()) - 1, nm = std::atoi(vt[r + 2].c_str()),
                      cen = std::atoi(vt[r + 3].c_str());
            r += 4;
            if (nm > 15 || r + nm > vt.size()) { err = "virtual: too many atoms in a virtual site"; return IBX_E_INVALID; }
            st.vs_type.push_back(ty);
            st.vs_center.push_back(cen - 1);
            for (int a = 0; a < nm; ++a) st.vs_atoms.push_back(std::atoi(vt[r + a].c_str()) - 1);
            r += nm;
            st.vs_off.push_back((int)st.vs_atoms.size());
        }
        std::vector<std::string> bl, al;
        if (read_lines(d + "virtbonds", bl))
            for (auto &l : bl) {
                auto tt = tokens(l);
                if (tt.size() < 2) break;
                st.virtbonds.push_back(std::atoi(tt[0].c_str()) - 1);
                st.virtbonds.push_back(std::atoi(tt[1].c_str()) - 1);
            }
        if (read_lines(d + "virtangles", al))
            for (auto &l : al) {
                auto tt = tokens(l);
                if (tt.size() < 3) break;
                for (int k = 0; k < 3; ++k) st.virtangles.push_back(std::atoi(tt[k].c_str()) - 1);
            }
    }

    D.natoms = natoms;
    D.nvs = nvs;
    D.ntype = ntype;
    D.mass = st.mass.data();
    D.label = st.label.data();
    D.itype = st.itype.data();
    D.nbonds = st.nbonds.data();
    D.jbond = st.jbond.data();
    D.s = st.s.data();
    D.v = st.v.data();
    ibx_tables *tabs[5] = {&D.bond, &D.angle, &D.tors, &D.nb, &D.oop};
    for (int k = 0; k < 5; ++k) {
        tabs[k]->ntab = (int)st.toff[k].size();
        tabs[k]->types = st.ttypes[k].data();
        tabs[k]->off = st.toff[k].data();
        tabs[k]->nrow = st.tnrow[k].data();
        tabs[k]->x = st.tx[k].data();
        tabs[k]->v = st.tv[k].data();
    }
    D.vs_type = st.vs_type.data();
    D.vs_center = st.vs_center.data();
    D.vs_off = st.vs_off.data();
    D.vs_atoms = st.vs_atoms.data();
    D.nvirtbonds = (int)st.virtbonds.size() / 2;
    D.virtbonds = st.virtbonds.data();
    D.nvirtangles = (int)st.virtangles.size() / 3;
    D.virtangles = st.virtangles.data();
    return 0;
}

}  // namespace ibx
