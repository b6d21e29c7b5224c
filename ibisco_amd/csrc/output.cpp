// Trajectory and restart output in the reference's formats (WRITETRJ.F90).
//
// A TRZ frame is a run of Fortran unformatted sequential records (each
// payload between two 4-byte length markers), as WRITE(113) produces with the
// file opened FORM='UNFORMATTED', ACCESS='SEQUENTIAL' (main.F90:96).  The
// restart file is the text form RDCOOR.f90 reads back as a coordinate file.
// Quantities arrive in the reduced units the reference holds (UNIT.f90) and
// are converted here, in FP64.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ibx.h"
#include "ibx_internal.h"

namespace {

// UNIT.f90:9-23 in FP64 (module_var.F90:14-17 constants)
const double kKB = 1.380658e-23, kNA = 6.0221367e23, kMMOLY = 0.012;
const double kMassScale = kMMOLY / kNA, kRScale = 1.0e-9, kEScale = kKB * 300.0;
const double kTimeScale = kRScale * std::sqrt(kMassScale / kEScale);   // reduced time -> s
const double kVScale = kRScale / kTimeScale;                            // reduced velocity -> m/s
const double kPScale = kEScale / (kRScale * kRScale * kRScale) / 1000.0;   // reduced pressure -> kPa
const double kConv = kNA * kEScale / 1000.0;                            // reduced energy -> kJ/mol

thread_local std::string g_out_err;

int oerr(int code, const std::string &msg) {
    g_out_err = msg;
    return code;
}

// One record: length, payload, length.
struct Rec {
    std::vector<unsigned char> b;
    template <typename T>
    Rec &put(T x) {
        const unsigned char *p = reinterpret_cast<const unsigned char *>(&x);
        b.insert(b.end(), p, p + sizeof(T));
        return *this;
    }
    bool write(FILE *f) const {
        const uint32_t n = (uint32_t)b.size();
        return fwrite(&n, 4, 1, f) == 1 && (n == 0 || fwrite(b.data(), 1, n, f) == n) && fwrite(&n, 4, 1, f) == 1;
    }
};

}  // namespace

extern "C" {

const char *ibx_output_last_error(void) { return g_out_err.c_str(); }

int ibx_trz_write_frame(const char *path, const char *title, int frame, int ntrj, int natoms, double time_ps,
                        const double box[3], const double pt[6], double epot, double ek, double mktemp,
                        const double *s, const double *v) {
    if (!path || frame < 1 || natoms < 0 || !box || !pt || (natoms > 0 && (!s || !v)))
        return oerr(IBX_E_INVALID, "invalid arguments");
    FILE *f = std::fopen(path, frame == 1 ? "wb" : "ab");
    if (!f) return oerr(IBX_E_IO, std::string("cannot open ") + path);
    bool ok = true;
    if (frame == 1) {   // WRITETRJ.F90:50-55: the title (CHARACTER(72) in CHARACTER(80)), then nrec = 10
        Rec t;
        char buf[80];
        std::memset(buf, ' ', sizeof buf);
        if (title) std::memcpy(buf, title, std::min<size_t>(std::strlen(title), 72));
        t.b.assign(buf, buf + 80);
        ok = ok && t.write(f) && Rec().put<int32_t>(10).write(f);
    }
    // frame number, its step index, atoms, time (WRITETRJ.F90:58)
    ok = ok && Rec().put<int32_t>(frame).put<int32_t>(ntrj * frame).put<int32_t>(natoms).put<double>(time_ps).write(f);
    {   // the cell as a 3x3 matrix (WRITETRJ.F90:64-66)
        Rec r;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) r.put<double>(a == b ? box[a] : 0.0);
        ok = ok && r.write(f);
    }
    {   // isotropic pressure and the tensor, kPa (WRITETRJ.F90:69-79): pt is 11 22 33 12 13 23
        Rec r;
        r.put<double>((pt[0] + pt[1] + pt[2]) * kPScale / 3.0).put<double>(pt[0] * kPScale).put<double>(pt[3] * kPScale);
        r.put<double>(pt[1] * kPScale).put<double>(pt[4] * kPScale).put<double>(pt[5] * kPScale);
        r.put<double>(pt[2] * kPScale);
        ok = ok && r.write(f);
    }
    // 6, total, potential and kinetic energy (kJ/mol), temperature, 0, 0 (WRITETRJ.F90:81-85); the
    // temperature is SUM(EK)*MKTEMP unscaled, in units of 300 K, as the reference writes it
    ok = ok && Rec().put<int32_t>(6).put<double>((epot + ek) * kConv).put<double>(epot * kConv).put<double>(ek * kConv)
                   .put<double>(ek * mktemp).put<double>(0.0).put<double>(0.0).write(f);
    // positions (nm) and velocities (nm/ps) in single precision, one record per component (WRITETRJ.F90:88-96)
    const double vconv = 1.0 / 1.0e12 / kTimeScale;
    for (int k = 0; k < 3 && ok; ++k) {
        Rec r;
        for (int i = 0; i < natoms; ++i) r.put<float>((float)s[3 * (size_t)i + k]);
        ok = r.write(f);
    }
    for (int k = 0; k < 3 && ok; ++k) {
        Rec r;
        for (int i = 0; i < natoms; ++i) r.put<float>((float)(v[3 * (size_t)i + k] * vconv));
        ok = r.write(f);
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? IBX_OK : oerr(IBX_E_IO, std::string("write failed: ") + path);
}

int ibx_restart_write(const char *path, const char *title, double time_ps, const double box[3], int natoms,
                      const int *itype, const int *nbonds, const int *jbond, const double *s, const double *v,
                      int nmol, const int *natm, const char *const *names) {
    if (!path || !box || natoms < 0 || nmol < 0 || (natoms > 0 && (!itype || !nbonds || !jbond || !s || !v)) ||
        (nmol > 0 && !natm))
        return oerr(IBX_E_INVALID, "invalid arguments");
    int64_t sum = 0;
    for (int m = 0; m < nmol; ++m) sum += natm[m];
    if (sum != natoms) return oerr(IBX_E_INVALID, "molecule sizes do not add up to the atoms");
    FILE *f = std::fopen(path, "w");
    if (!f) return oerr(IBX_E_IO, std::string("cannot open ") + path);
    // WRITETRJ.F90:131-141: title, time, box, the record description
    std::fprintf(f, " %s\n", title ? title : "");
    std::fprintf(f, " Time\t %.9g\t(ps)\n", time_ps);
    std::fprintf(f, " ***** Box Length(X, Y, Z in nanometres) **\n");
    std::fprintf(f, " %.17g %.17g %.17g\n", box[0], box[1], box[2]);
    const char *stars70 = "**********************************************************************";
    std::fprintf(f, "%s\n", stars70);
    std::fprintf(f, "********** Record for each atom is in the form:-            **********\n");
    std::fprintf(f, "********** Index Atom_Type No._of_bonds X Y Z (coords.in m) **********\n");
    std::fprintf(f, "********** Vx Vy Vz (in m/s) Indices_of_bonded_atoms        **********\n");
    std::fprintf(f, "%s\n", stars70);
    std::fprintf(f, " num_of_molecules   %d\n", nmol);
    // per molecule its size and name, per atom (9040) index, type, bonds, position, then (9050-9090) the
    // velocity in nm/ps and the bonded atoms; indices 1-based
    const double vconv = kVScale * 1.0e-3;
    int l = 0;
    for (int m = 0; m < nmol; ++m) {
        std::fprintf(f, " %d  Atoms_in_Molecule_No. %d %s\n", natm[m], m + 1, names && names[m] ? names[m] : "mol");
        for (int a = 0; a < natm[m]; ++a, ++l) {
            std::fprintf(f, " %6d %2d %1d %21.13E %21.13E %21.13E\n", l + 1, itype[l] + 1, nbonds[l], s[3 * (size_t)l],
                         s[3 * (size_t)l + 1], s[3 * (size_t)l + 2]);
            std::fprintf(f, "%16.10f %16.10f %16.10f", v[3 * (size_t)l] * vconv, v[3 * (size_t)l + 1] * vconv,
                         v[3 * (size_t)l + 2] * vconv);
            for (int b = 0; b < nbonds[l] && b < 4; ++b) std::fprintf(f, " %5d", jbond[4 * (size_t)l + b] + 1);
            std::fprintf(f, "\n");
        }
    }
    return std::fclose(f) == 0 ? IBX_OK : oerr(IBX_E_IO, std::string("write failed: ") + path);
}

int ibx_write_trz(ibx_ctx *ctx, int sys, const char *path, const char *title, int frame, int ntrj, int step,
                  double initime_ps) {
    const ibx::HostSystem *h = ibx::host_system(ctx, sys);
    if (!h) return oerr(IBX_E_INVALID, "invalid context or system index");
    std::vector<double> s(3 * (size_t)h->nitems), v(3 * (size_t)h->natoms);
    double obs[32], box[3];
    int rc = ibx_get_state(ctx, sys, s.data(), v.data());
    if (!rc) rc = ibx_get_observables(ctx, sys, obs);
    if (!rc) rc = ibx_get_box(ctx, sys, box);   // NPT: the box of the last step
    if (rc) return oerr(rc, ibx_last_error(ctx));
    // obs: 0..14 V_NB, V_BOND, V_ANGLE, V_TORSION, V_OOP (3 classes each), 15..20 the pressure tensor, 21..22
    // the kinetic energy of each population
    double epot = 0.0;
    for (int k = 0; k < 15; ++k) epot += obs[k];
    const double pt[6] = {obs[15], obs[16], obs[17], obs[18], obs[19], obs[20]};
    const double time_ps = (double)step * h->dt * kTimeScale * 1.0e12 + initime_ps;   // WRITETRJ.F90:57
    return ibx_trz_write_frame(path, title, frame, ntrj, h->natoms, time_ps, box, pt, epot, obs[21] + obs[22],
                               h->mktemp, s.data(), v.data());
}

int ibx_write_restart(ibx_ctx *ctx, int sys, const char *path, const char *title, int step, double initime_ps,
                      int nmol, const int *natm, const char *const *names) {
    const ibx::HostSystem *h = ibx::host_system(ctx, sys);
    if (!h) return oerr(IBX_E_INVALID, "invalid context or system index");
    std::vector<double> s(3 * (size_t)h->nitems), v(3 * (size_t)h->natoms);
    double box[3];
    int rc = ibx_get_state(ctx, sys, s.data(), v.data());
    if (!rc) rc = ibx_get_box(ctx, sys, box);
    if (rc) return oerr(rc, ibx_last_error(ctx));
    const double time_ps = (double)step * h->dt * kTimeScale * 1.0e12 + initime_ps;
    return ibx_restart_write(path, title, time_ps, box, h->natoms, h->itype.data(), h->nbonds.data(),
                             h->jbond.data(), s.data(), v.data(), nmol, natm, names);
}

}  // extern "C"
