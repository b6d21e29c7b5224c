"""Particle systems in IBIsCO's input formats, and seeded synthetic generators.

A :class:`System` holds what the reference reads from its input directory:

* ``control``      keyword file (``RDCONTROL.f90:45-536``),
* ``interaction``  type table plus one 2-column table file per interaction
  (``RDINTERACT.f90:56-537``),
* ``coordinate``   box, molecules, per-atom type, bonds, positions (nm) and
  velocities (nm/ps) (``RDCOOR.f90:35-139``),
* ``virtual``, ``virtbonds``, ``virtangles`` for hybrid models
  (``RDVIRTUAL.f90``, ``RDVIRTBONDS.f90:42-64``, ``RDVIRTANGLES.f90:36-50``).

Indices are 0-based here and 1-based in the files.  :func:`write_dir` and
:func:`read_dir` convert both ways; the generators build the synthetic
melts the benchmarks run on (coarse-grained and united-atom polymer melts,
Gaussian-corrected tables for the replica and serving workloads).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import units

MAXBONDS = 4  # ALLOCATEVAR.f90:13


@dataclass
class Table:
    """A tabulated potential: abscissa (nm or degrees) and energy (kJ/mol)."""
    types: tuple
    x: np.ndarray
    v: np.ndarray
    name: str = ""


@dataclass
class Control:
    ensemble: str = "NVE"
    temperature: float = 300.0        # K
    pressure: float = 101.3           # kPa
    nstep: int = 100
    dt: float = 0.005                 # ps
    cutoff: float = 1.5               # nm
    rlist: float = 1.8                # nm
    bead_cutoff: float = 1.5
    bead_rlist: float = 1.8
    nupdate: int = 10
    taut: float = 0.1                 # ps
    taup: float = 5.0
    beta: float = 1.0e-6
    trajectory: int = 100
    halt_drift: int = 1000000
    non_bonded: int = 4
    non_bonded_bead: int = 4
    init_velocities: bool = False
    mts: int = 0


@dataclass
class System:
    box: np.ndarray
    type_names: list
    type_mass: np.ndarray            # kg/mol
    type_label: list                 # 'A' atom or 'B' bead, per type
    itype: np.ndarray                # [natoms] 0-based type of each atom
    s: np.ndarray                    # [natoms,3] unwrapped positions, nm
    v: np.ndarray                    # [natoms,3] velocities, nm/ps
    jbond: list                      # per atom: list of bonded partners (0-based), file order
    mol_sizes: list
    mol_names: list
    bond_tables: list = field(default_factory=list)
    angle_tables: list = field(default_factory=list)
    torsion_tables: list = field(default_factory=list)
    nb_tables: list = field(default_factory=list)
    oop_tables: list = field(default_factory=list)
    control: Control = field(default_factory=Control)
    # hybrid model: virtual sites appended after the atoms
    vs_type: np.ndarray = None       # [nvs] 0-based type
    vs_center: np.ndarray = None     # [nvs] centre atom (0-based) or -1 for a centre of mass
    vs_atoms: list = None            # [nvs] arrays of member atoms (0-based)
    virtbonds: np.ndarray = None     # [k,2] global 0-based
    virtangles: np.ndarray = None    # [k,3] global 0-based
    time_ps: float = 0.0
    title: str = "synthetic"

    @property
    def natoms(self):
        return int(self.itype.shape[0])

    @property
    def nvs(self):
        return 0 if self.vs_type is None else int(len(self.vs_type))

    def bonds(self):
        """Unique bonds (i<j) as an int32 array [nb,2]."""
        out = [(i, j) for i, js in enumerate(self.jbond) for j in js if j > i]
        return np.array(out, dtype=np.int32).reshape(-1, 2)


# ----------------------------------------------------------------------------------------------
# tables


def lj_sf_table(sigma, eps, rc, dr, cap_frac=0.6):
    """Shifted-force LJ 12-6 (kJ/mol) sampled at r = k*dr, k = 0..rc/dr.

    Below ``cap_frac*sigma`` the potential continues linearly (constant force)
    so that the F/r table stays finite in FP32.
    """
    n = int(round(rc / dr)) + 1
    r = np.arange(n, dtype=np.float64) * dr

    def v_lj(x):
        sr6 = (sigma / x) ** 6
        return 4.0 * eps * (sr6 * sr6 - sr6)

    def dv_lj(x):
        sr6 = (sigma / x) ** 6
        return 4.0 * eps * (-12.0 * sr6 * sr6 + 6.0 * sr6) / x

    vc, dvc = v_lj(rc), dv_lj(rc)
    rcap = cap_frac * sigma
    rr = np.maximum(r, rcap)
    v = v_lj(rr) - vc - (rr - rc) * dvc
    below = r < rcap
    slope = dv_lj(rcap) - dvc
    v[below] = v[below] + (r[below] - rcap) * slope
    return r, v


def gaussian_correction(r, rng, ngauss=8, amp=0.5, wmin=0.05, wmax=0.2, mu_lo=0.4, mu_hi=1.4):
    """Sum of seeded Gaussians, re-shifted so that V(rc)=0 and V'(rc)=0 still hold."""
    a = rng.uniform(-amp, amp, ngauss)
    w = rng.uniform(wmin, wmax, ngauss)
    mu = rng.uniform(mu_lo, mu_hi, ngauss)
    g = np.zeros_like(r)
    dg_rc = 0.0
    g_rc = 0.0
    rc = r[-1]
    for k in range(ngauss):
        g += a[k] * np.exp(-((r - mu[k]) ** 2) / (2 * w[k] ** 2))
        e = a[k] * math.exp(-((rc - mu[k]) ** 2) / (2 * w[k] ** 2))
        g_rc += e
        dg_rc += e * (-(rc - mu[k]) / w[k] ** 2)
    return g - g_rc - (r - rc) * dg_rc


def harmonic_bond_table(r0, k, rmax, dr):
    n = int(round(rmax / dr)) + 1
    r = np.arange(n, dtype=np.float64) * dr
    return r, 0.5 * k * (r - r0) ** 2


def harmonic_angle_table(theta0_deg, k_rad, dtheta=0.5):
    n = int(round(180.0 / dtheta)) + 1
    th = np.arange(n, dtype=np.float64) * dtheta
    d = (th - theta0_deg) * units.D2R
    return th, 0.5 * k_rad * d * d


RB_ALKANE = (9.28, 12.16, -13.12, -3.06, 26.24, -31.5)  # kJ/mol, Ryckaert-Bellemans


def rb_torsion_table(coeffs=RB_ALKANE, dphi=0.5):
    """Ryckaert-Bellemans torsion over phi in [0,360] with trans at 180 degrees."""
    n = int(round(360.0 / dphi)) + 1
    phi = np.arange(n, dtype=np.float64) * dphi
    c = np.cos((phi - 180.0) * units.D2R)
    v = np.zeros(n)
    for p, cn in enumerate(coeffs):
        v += cn * c ** p
    return phi, v


def oop_table(k_kjmol=20.0, dphi=2.0):
    """Out-of-plane potential K (1 - cos phi) over the signed angle phi in [-180, 180] degrees."""
    n = int(round(360.0 / dphi)) + 1
    phi = -180.0 + np.arange(n, dtype=np.float64) * dphi
    return phi, k_kjmol * (1.0 - np.cos(phi * units.D2R))


# ----------------------------------------------------------------------------------------------
# generators


def _grow_chains(rng, nchains, nbeads, box, b, theta0_deg, spread_deg=8.0):
    """Random-walk chains with fixed bond length and bond angle near theta0 (unwrapped nm)."""
    pos = np.empty((nchains, nbeads, 3))
    pos[:, 0] = rng.uniform(-0.5, 0.5, (nchains, 3)) * box
    u = rng.normal(size=(nchains, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    if nbeads > 1:
        pos[:, 1] = pos[:, 0] + b * u
    for k in range(2, nbeads):
        turn = np.deg2rad(180.0 - theta0_deg + rng.uniform(-spread_deg, spread_deg, nchains))
        # perpendicular direction with random azimuth
        a = rng.normal(size=(nchains, 3))
        a -= (a * u).sum(1, keepdims=True) * u
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        u = np.cos(turn)[:, None] * u + np.sin(turn)[:, None] * a
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        pos[:, k] = pos[:, k - 1] + b * u
    return pos.reshape(-1, 3)


def _grow_chains_avoiding(rng, nchains, nbeads, box, b, theta0_deg, dmin, spread_deg=8.0, tries=40):
    """Random-walk chains whose beads keep dmin from every earlier bead (bonded
    neighbours and 1-3 partners aside); a spatial hash keeps it O(N).  Meant for
    the small systems of the tests; growth falls back to the best of its tries."""
    L = np.asarray(box, dtype=np.float64)
    nc = np.maximum(1, np.floor(L / dmin).astype(int))
    cs = L / nc
    grid = {}
    out = np.empty((nchains * nbeads, 3))

    def key(p):
        return tuple((np.floor((p / L + 0.5) % 1.0 * nc).astype(int)) % nc)

    def min_dist(p, skip):
        k = key(p)
        best = 1e9
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    cell = ((k[0] + dx) % nc[0], (k[1] + dy) % nc[1], (k[2] + dz) % nc[2])
                    for q in grid.get(cell, ()):
                        if q in skip:
                            continue
                        d = out[q] - p
                        d -= np.round(d / L) * L
                        best = min(best, float(np.sqrt((d * d).sum())))
        return best

    def add(idx):
        grid.setdefault(key(out[idx]), []).append(idx)

    n = 0
    for _c in range(nchains):
        first = n
        best_p, best_d = None, -1.0
        for _t in range(tries):
            p = rng.uniform(-0.5, 0.5, 3) * L
            dd = min_dist(p, ())
            if dd > best_d:
                best_p, best_d = p, dd
            if dd >= dmin:
                break
        out[n] = best_p
        add(n)
        n += 1
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        for k in range(1, nbeads):
            best_p, best_d, best_u = None, -1.0, None
            for _t in range(tries):
                if k == 1:
                    un = rng.normal(size=3)
                else:
                    turn = np.deg2rad(180.0 - theta0_deg + rng.uniform(-spread_deg, spread_deg))
                    a = rng.normal(size=3)
                    a -= (a @ u) * u
                    a /= np.linalg.norm(a)
                    un = np.cos(turn) * u + np.sin(turn) * a
                un = un / np.linalg.norm(un)
                p = out[n - 1] + b * un
                dd = min_dist(p, {n - 1, n - 2} if k > 1 else {n - 1})
                if dd > best_d:
                    best_p, best_d, best_u = p, dd, un
                if dd >= dmin:
                    break
            out[n] = best_p
            u = best_u
            add(n)
            n += 1
        del first
    return out


def maxwell_velocities(rng, mass_kgmol, temp_k):
    """Per-component Gaussian velocities (nm/ps) at temp_k with zero total momentum."""
    m = np.asarray(mass_kgmol, dtype=np.float64)
    sig = np.sqrt(units.KB * temp_k / (m / units.NA)) / 1.0e3
    v = rng.normal(size=(m.shape[0], 3)) * sig[:, None]
    p = (v * m[:, None]).sum(0) / m.sum()
    return v - p[None, :]


def polymer_melt(nchains, nbeads, density, mass_kgmol, temp_k, seed, *, sigma, eps, rc, dr_nb,
                 bond_r0, bond_k, angle_theta0, angle_k, bond_rmax=None, bond_dr=None,
                 torsion=False, label="B", nb_table=None, dt=0.005, rlist=None, nupdate=10,
                 nstep=1000, non_bonded=4, name="melt", avoid=None, box=None):
    """Linear homopolymer melt of one particle type in a periodic cube.

    ``density`` is in particles per nm^3.  ``avoid`` (a fraction of sigma) grows
    the chains with a minimum distance between non-bonded beads; without it the
    walks overlap freely.  ``box`` (three edges, nm) replaces the cube of the
    density's volume by an orthorhombic box.  Bonds and angles are harmonic
    (V = k/2 (x-x0)^2, k in kJ/mol/nm^2 and kJ/mol/rad^2), the nonbonded
    table is shifted-force LJ unless ``nb_table`` (r, V) is given.
    """
    rng = np.random.default_rng(seed)
    n = nchains * nbeads
    if box is None:
        L = (n / density) ** (1.0 / 3.0)
        box = np.array([L, L, L])
    else:
        box = np.asarray(box, dtype=np.float64)
    if avoid:
        s = _grow_chains_avoiding(rng, nchains, nbeads, box, bond_r0, angle_theta0, avoid * sigma)
    else:
        s = _grow_chains(rng, nchains, nbeads, box, bond_r0, angle_theta0)
    itype = np.zeros(n, dtype=np.int32)
    v = maxwell_velocities(rng, np.full(n, mass_kgmol), temp_k)
    jbond = []
    for c in range(nchains):
        for k in range(nbeads):
            lst = []
            if k > 0:
                lst.append(c * nbeads + k - 1)
            if k < nbeads - 1:
                lst.append(c * nbeads + k + 1)
            jbond.append(lst)
    if nb_table is None:
        nb_table = lj_sf_table(sigma, eps, rc, dr_nb)
    bond_rmax = bond_rmax if bond_rmax is not None else 2.5 * bond_r0
    bond_dr = bond_dr if bond_dr is not None else bond_r0 / 400.0
    br, bv = harmonic_bond_table(bond_r0, bond_k, bond_rmax, bond_dr)
    at, av = harmonic_angle_table(angle_theta0, angle_k)
    sysm = System(box=box, type_names=["C"], type_mass=np.array([mass_kgmol]), type_label=[label],
                  itype=itype, s=s, v=v, jbond=jbond, mol_sizes=[nbeads] * nchains,
                  mol_names=[name] * nchains)
    sysm.bond_tables = [Table((0, 0), br, bv, "bond1")]
    sysm.angle_tables = [Table((0, 0, 0), at, av, "ang1")]
    if torsion:
        pt, pv = rb_torsion_table()
        sysm.torsion_tables = [Table((0, 0, 0, 0), pt, pv, "tors1")]
    sysm.nb_tables = [Table((0, 0), nb_table[0], nb_table[1], "nb1")]
    sysm.control = Control(ensemble="NVE", temperature=temp_k, nstep=nstep, dt=dt, cutoff=rc,
                           rlist=rlist if rlist is not None else rc + 0.3, nupdate=nupdate,
                           non_bonded=non_bonded)
    return sysm


def hybrid_melt(n_at_chains, n_cg_chains, chain_len=60, group=6, seed=0, temp_k=450.0, rc=1.4, skin=0.3,
                dr_nb=0.002, dt=0.001, nupdate=10, atom_density=33.5, avoid=None, light=False):
    """Hybrid atomistic/CG melt: atomistic chains of ``chain_len`` CH2-like atoms
    (type A) carrying one centre-of-mass virtual site (type V) per ``group``
    consecutive atoms, and CG chains of chain_len/group beads (type B, each the
    mass of ``group`` atoms).  Three LJ tables with sigma drawn from 0.35-0.50 nm
    and epsilon from 0.2-3.0 kJ/mol (atom-atom, bead-bead, bead-site; the
    site-site table is the bead-bead one and never used, as virtual-site pairs
    are always excluded), shifted force to rc at dr_nb.  The box holds the atoms
    at ``atom_density`` per nm^3, a bead taking the volume of ``group`` atoms.
    With ``light`` every second atom of an atomistic chain is a light type H
    (1.008 g/mol, the same tables as A), so centres of mass differ from means."""
    rng = np.random.default_rng(seed)
    m_atom = 0.014027
    nb_cg = chain_len // group
    n_at = n_at_chains * chain_len
    n_cg = n_cg_chains * nb_cg
    L = ((n_at + group * n_cg) / atom_density) ** (1.0 / 3.0)
    box = np.array([L, L, L])
    if avoid:
        sa = _grow_chains_avoiding(rng, n_at_chains, chain_len, box, 0.153, 112.0, avoid * 0.35)
        sb = _grow_chains_avoiding(rng, n_cg_chains, nb_cg, box, 0.5, 130.0, avoid * 0.35)
    else:
        sa = _grow_chains(rng, n_at_chains, chain_len, box, 0.153, 112.0)
        sb = _grow_chains(rng, n_cg_chains, nb_cg, box, 0.5, 130.0)
    s = np.concatenate([sa.reshape(-1, 3), sb.reshape(-1, 3)])
    itype = np.concatenate([np.zeros(n_at, np.int32), np.ones(n_cg, np.int32)])
    masses = np.array([m_atom, group * m_atom, group * m_atom, 0.001008])
    if light:
        itype[1:n_at:2] = 3
    v = maxwell_velocities(rng, masses[itype], temp_k)
    jbond = []
    for nch, clen, base in ((n_at_chains, chain_len, 0), (n_cg_chains, nb_cg, n_at)):
        for c in range(nch):
            for k in range(clen):
                i = base + c * clen + k
                jbond.append(([i - 1] if k > 0 else []) + ([i + 1] if k < clen - 1 else []))
    vs_atoms = [np.arange(c * chain_len + g * group, c * chain_len + (g + 1) * group, dtype=np.int32)
                for c in range(n_at_chains) for g in range(chain_len // group)]
    pars = [(rng.uniform(0.35, 0.50), rng.uniform(0.2, 3.0)) for _ in range(3)]
    tabs = [lj_sf_table(sg, ep, rc, dr_nb) for sg, ep in pars]
    names, labels = ["A", "B", "V", "H"], ["A", "B", "B", "A"]
    if not light:
        names, labels, masses = names[:3], labels[:3], masses[:3]
    sysm = System(box=box, type_names=names, type_mass=masses, type_label=labels,
                  itype=itype, s=s, v=v, jbond=jbond,
                  mol_sizes=[chain_len] * n_at_chains + [nb_cg] * n_cg_chains,
                  mol_names=["atomistic"] * n_at_chains + ["coarse"] * n_cg_chains)
    br, bv = harmonic_bond_table(0.153, 250000.0, 0.4, 0.0005)
    cr, cv = harmonic_bond_table(0.5, 5000.0, 1.3, 0.001)
    at, av = harmonic_angle_table(112.0, 520.0)
    ct, cvv = harmonic_angle_table(130.0, 25.0)
    sysm.bond_tables = [Table((0, 0), br, bv, "bondA"), Table((1, 1), cr, cv, "bondB")]
    sysm.angle_tables = [Table((0, 0, 0), at, av, "angA"), Table((1, 1, 1), ct, cvv, "angB")]
    sysm.nb_tables = [Table((0, 0), tabs[0][0], tabs[0][1], "nbAA"), Table((1, 1), tabs[1][0], tabs[1][1], "nbBB"),
                      Table((1, 2), tabs[2][0], tabs[2][1], "nbBV"), Table((2, 2), tabs[1][0], tabs[1][1], "nbVV")]
    if light:
        sysm.bond_tables[0] = Table((0, 3), br, bv, "bondAH")
        sysm.angle_tables[0] = Table((0, 3, 0), at, av, "angAHA")
        sysm.angle_tables.append(Table((3, 0, 3), at, av, "angHAH"))
        sysm.nb_tables += [Table((0, 3), tabs[0][0], tabs[0][1], "nbAH"), Table((3, 3), tabs[0][0], tabs[0][1], "nbHH")]
    sysm.vs_type = np.full(len(vs_atoms), 2, np.int32)
    sysm.vs_center = np.full(len(vs_atoms), -1, np.int32)
    sysm.vs_atoms = vs_atoms
    sysm.control = Control(ensemble="NVE", temperature=temp_k, nstep=1000, dt=dt, cutoff=rc, rlist=rc + skin,
                           bead_cutoff=rc, bead_rlist=rc + skin, nupdate=nupdate, non_bonded=4, non_bonded_bead=4)
    return sysm


def comb_melt(nchains=20, nbackbone=40, seed=0, temp_k=400.0, rc=1.0, rlist=1.3, dt=0.001, nupdate=10,
              site_density=33.0, spread_deg=25.0, oop_only=False):
    """Comb polymer melt with out-of-plane terms: backbone atoms C (type 0) carry a
    side atom S (type 1) on every odd backbone position, so those atoms have three
    bonds (previous, next, side, in that file order) and an improper C-C-C-S
    (OUTOFPLANES.inc) with a K (1 - cos phi) table over [-180, 180] degrees.
    Side atoms leave the backbone plane by up to ``spread_deg``.  Harmonic bonds
    and angles, LJ between all sites, 1-2 and 1-3 exclusions.  ``oop_only`` zeroes
    every table but the out-of-plane one (the exclusions and terms stay), so the
    forces are the out-of-plane forces alone."""
    rng = np.random.default_rng(seed)
    nside = len(range(1, nbackbone - 1, 2))
    per = nbackbone + nside
    n = nchains * per
    L = (n / site_density) ** (1.0 / 3.0)
    box = np.array([L, L, L])
    bb = _grow_chains_avoiding(rng, nchains, nbackbone, box, 0.153, 112.0, 0.3).reshape(nchains, nbackbone, 3)
    s = np.empty((nchains, per, 3))
    itype = np.zeros(n, np.int32)
    jbond = [None] * n
    for c in range(nchains):
        base = c * per
        s[c, :nbackbone] = bb[c]
        for k in range(nbackbone):
            jbond[base + k] = ([base + k - 1] if k > 0 else []) + ([base + k + 1] if k < nbackbone - 1 else [])
        for q, k in enumerate(range(1, nbackbone - 1, 2)):
            i = base + nbackbone + q
            a, b = bb[c, k - 1] - bb[c, k], bb[c, k + 1] - bb[c, k]
            out = -(a / np.linalg.norm(a) + b / np.linalg.norm(b))
            out /= np.linalg.norm(out)
            nrm = np.cross(a, b)
            nrm /= np.linalg.norm(nrm)
            tilt = np.deg2rad(rng.uniform(-spread_deg, spread_deg))
            d = np.cos(tilt) * out + np.sin(tilt) * nrm
            s[c, nbackbone + q] = bb[c, k] + 0.153 * d
            itype[i] = 1
            jbond[base + k].append(i)
            jbond[i] = [base + k]
    s = s.reshape(-1, 3)
    masses = np.array([0.013019, 0.015035])
    v = maxwell_velocities(rng, masses[itype], temp_k)
    sysm = System(box=box, type_names=["C", "S"], type_mass=masses, type_label=["A", "A"], itype=itype, s=s,
                  v=v, jbond=jbond, mol_sizes=[per] * nchains, mol_names=["comb"] * nchains)
    br, bv = harmonic_bond_table(0.153, 250000.0, 0.4, 0.0005)
    at, av = harmonic_angle_table(112.0, 520.0)
    at2, av2 = harmonic_angle_table(110.0, 400.0)
    pt, pv = oop_table()
    nbr, nbv = lj_sf_table(0.38, 0.45, rc, 0.002)
    sysm.bond_tables = [Table((0, 0), br, bv, "bondCC"), Table((0, 1), br, bv, "bondCS")]
    sysm.angle_tables = [Table((0, 0, 0), at, av, "angCCC"), Table((0, 0, 1), at2, av2, "angCCS")]
    sysm.oop_tables = [Table((0, 0, 0, 1), pt, pv, "oopCCCS")]
    sysm.nb_tables = [Table((0, 0), nbr, nbv, "nbCC"), Table((0, 1), nbr, nbv, "nbCS"),
                      Table((1, 1), nbr, nbv, "nbSS")]
    if oop_only:
        for tb in sysm.bond_tables + sysm.angle_tables + sysm.nb_tables:
            tb.v = np.zeros_like(tb.v)
    sysm.control = Control(ensemble="NVE", temperature=temp_k, nstep=100, dt=dt, cutoff=rc, rlist=rlist,
                           nupdate=nupdate, non_bonded=4)
    return sysm


def cg_melt(nchains=200, nbeads=50, seed=0, density=8.0, temp_k=500.0, nb_table=None, **kw):
    """Coarse-grained melt: 72 amu beads, LJ(0.47 nm, 3.5 kJ/mol) to 1.5 nm at 0.002 nm."""
    args = dict(mass_kgmol=0.072, temp_k=temp_k, seed=seed, sigma=0.47, eps=3.5, rc=1.5,
                dr_nb=0.002, bond_r0=0.47, bond_k=5000.0, angle_theta0=130.0, angle_k=25.0,
                bond_rmax=1.2, bond_dr=0.001, dt=0.005, rlist=1.8, nb_table=nb_table)
    args.update(kw)
    return polymer_melt(nchains, nbeads, density, **args)


def ua_melt(nchains=400, nbeads=250, seed=1, temp_k=450.0, **kw):
    """United-atom polyethylene-like melt: 14 amu CH2, 0.78 g/cm^3, LJ(0.401, 0.39) to 1.0 nm."""
    mass = 0.014027
    density = 0.78e3 / (mass / units.NA) * 1e-27  # sites per nm^3
    args = dict(mass_kgmol=mass, temp_k=temp_k, seed=seed, sigma=0.401, eps=0.39, rc=1.0,
                dr_nb=0.002, bond_r0=0.153, bond_k=250000.0, angle_theta0=112.0, angle_k=520.0,
                bond_rmax=0.4, bond_dr=0.0005, torsion=True, dt=0.001, rlist=1.3)
    args.update(kw)
    return polymer_melt(nchains, nbeads, density, **args)


def corrected_cg_table(seed, rc=1.5, dr=0.002):
    """The CG melt's LJ (0.47 nm, 3.5 kJ/mol) plus a seeded sum of 8 Gaussians (|A|<=0.5 kJ/mol, w in 0.05-0.2 nm)."""
    r, v = lj_sf_table(0.47, 3.5, rc, dr)
    rng = np.random.default_rng(10_000 + seed)
    return r, v + gaussian_correction(r, rng)


def replica_sizes(nrep, seed, nmin=2000, nmax=50000, cmin=20, cmax=100):
    """Log-uniform bead counts rounded to a multiple of a uniform chain length."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nrep):
        n = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
        c = int(rng.integers(cmin, cmax + 1))
        nch = max(1, int(round(n / c)))
        out.append((nch, c))
    return out


# ----------------------------------------------------------------------------------------------
# reference file formats


def _write_table(path, x, v):
    with open(path, "w") as f:
        for a, b in zip(x, v):
            f.write(f"{a:.10f} {b:.15e}\n")


def write_dir(sysm: System, path: str):
    """Write control, interaction (+ tables), coordinate and hybrid files in the reference formats."""
    os.makedirs(path, exist_ok=True)
    c = sysm.control
    hybrid = sysm.nvs > 0
    with open(os.path.join(path, "control"), "w") as f:
        kv = [
            ("ensemble", c.ensemble), ("temperature", c.temperature), ("pressure", c.pressure),
            ("atoms", sysm.natoms), ("num_of_time_steps", c.nstep), ("time_step", c.dt),
            ("temperature_coupling_time", c.taut), ("pressure_coupling_time", c.taup),
            ("isothermal_compressibility", c.beta), ("virtual_sites", sysm.nvs),
            ("bead_cutoff", c.bead_cutoff), ("bead_neighbour_list_cutoff", c.bead_rlist),
            ("non_bonded_bead", c.non_bonded_bead), ("cutoff", c.cutoff),
            ("neighbour_list_cutoff", c.rlist), ("update_neighbour_list", c.nupdate),
            ("trajectory", c.trajectory), ("halt_drift", c.halt_drift),
            ("non_bonded", c.non_bonded),
            ("initialize_velocities", "Y" if c.init_velocities else "N"), ("mts", c.mts),
        ]
        for k, val in kv:
            f.write(f"{k:<28s} {val}\n")
        f.write("end\n")
    with open(os.path.join(path, "interaction"), "w") as f:
        f.write(f"atom_types {len(sysm.type_names)}\n# type label mass A/B\n")
        for t, (nm, m, lab) in enumerate(zip(sysm.type_names, sysm.type_mass, sysm.type_label)):
            f.write(f"  {t + 1}  {nm}  {m:.10f}  {lab}\n")

        def section(title, tabs, ntyp, ext):
            f.write(f"{title} {len(tabs)}\n# types filename\n")
            for k, tb in enumerate(tabs):
                fn = tb.name or f"{ext}{k + 1}"
                f.write("  " + "  ".join(str(t + 1) for t in tb.types[:ntyp]) + f"  {fn}\n")
                _write_table(os.path.join(path, fn), tb.x, tb.v)

        section("bond_types", sysm.bond_tables, 2, "bond")
        section("angle_types", sysm.angle_tables, 3, "ang")
        section("torsion_types", sysm.torsion_tables, 4, "tors")
        section("non_bonded_interaction_types", sysm.nb_tables, 2, "nb")
        section("Out_of_Planes", sysm.oop_tables, 4, "oop")
    with open(os.path.join(path, "coordinate"), "w") as f:
        f.write(f" {sysm.title}\n")
        f.write(f" Time {sysm.time_ps:.10f} (ps)\n")
        f.write(" ********** Box Length(X, Y, Z in nanometres) **********\n")
        f.write(f"  {sysm.box[0]:.16f}  {sysm.box[1]:.16f}  {sysm.box[2]:.16f}\n")
        for _ in range(5):
            f.write(" **********************************************************************\n")
        f.write(f" num_of_molecules: {len(sysm.mol_sizes)}\n")
        a = 0
        lines = []
        for mi, (ms, mn) in enumerate(zip(sysm.mol_sizes, sysm.mol_names)):
            lines.append(f" {ms} atoms_in_molecule_no. {mi + 1} {mn}\n")
            for _ in range(ms):
                js = sysm.jbond[a]
                x, y, z = sysm.s[a]
                vx, vy, vz = sysm.v[a]
                lines.append(f" {a + 1} {sysm.itype[a] + 1} {len(js)} {x:.16e} {y:.16e} {z:.16e}\n")
                lines.append(f"   {vx:.16e} {vy:.16e} {vz:.16e} " + " ".join(str(j + 1) for j in js) + "\n")
                a += 1
        f.writelines(lines)
    if hybrid:
        with open(os.path.join(path, "virtual"), "w") as f:
            f.write("# virtual sites\n# count\n")
            f.write(f"{sysm.nvs}\n# index type natoms centre\n# member atoms\n")
            for k in range(sysm.nvs):
                mem = sysm.vs_atoms[k]
                f.write(f"{k + 1} {sysm.vs_type[k] + 1} {len(mem)} {sysm.vs_center[k] + 1}\n")
                f.write(" ".join(str(m + 1) for m in mem) + "\n")
        if sysm.virtbonds is not None and len(sysm.virtbonds):
            with open(os.path.join(path, "virtbonds"), "w") as f:
                for i, j in sysm.virtbonds:
                    f.write(f"{i + 1} {j + 1}\n")
        if sysm.virtangles is not None and len(sysm.virtangles):
            with open(os.path.join(path, "virtangles"), "w") as f:
                for i, j, k in sysm.virtangles:
                    f.write(f"{i + 1} {j + 1} {k + 1} 1\n")


def _tokens(line):
    """Whitespace tokens before a '#' comment (the PARSE.f90 tokenizer's behaviour on these files)."""
    return line.split("#", 1)[0].split()


def _fnum(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def _read_table(path):
    xs, vs = [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) < 2:
                break
            try:
                xs.append(_fnum(t[0]))
                vs.append(_fnum(t[1]))
            except ValueError:
                break
    return np.array(xs), np.array(vs)


def read_control(path):
    c = Control()
    vals = {}
    with open(path) as f:
        for line in f:
            t = _tokens(line)
            if len(t) >= 2 and t[0] not in vals:
                vals[t[0]] = t[1]
            elif len(t) >= 2 and t[0] == "mts":
                vals["mts"] = t[1]       # the last mts line wins (RDCONTROL.f90:488-491)
    c.ensemble = vals.get("ensemble", "NVE").upper()
    c.temperature = _fnum(vals.get("temperature", "300"))
    c.pressure = _fnum(vals.get("pressure", "101.3"))
    c.nstep = int(vals.get("num_of_time_steps", "0"))
    c.dt = _fnum(vals.get("time_step", "0.001"))
    c.cutoff = _fnum(vals.get("cutoff", "1.0"))
    c.rlist = _fnum(vals.get("neighbour_list_cutoff", "1.2"))
    c.bead_cutoff = _fnum(vals.get("bead_cutoff", str(c.cutoff)))
    c.bead_rlist = _fnum(vals.get("bead_neighbour_list_cutoff", str(c.rlist)))
    c.nupdate = int(vals.get("update_neighbour_list", "10"))
    c.taut = _fnum(vals.get("temperature_coupling_time", "0.1"))
    c.taup = _fnum(vals.get("pressure_coupling_time", "5.0"))
    c.beta = _fnum(vals.get("isothermal_compressibility", "1e-6"))
    c.trajectory = int(vals.get("trajectory", "100"))
    c.halt_drift = int(vals.get("halt_drift", "1000000"))
    c.non_bonded = int(vals.get("non_bonded", "4"))
    c.non_bonded_bead = int(vals.get("non_bonded_bead", "4"))
    c.init_velocities = vals.get("initialize_velocities", "N")[:1] in ("Y", "y")
    m = vals.get("mts", "0")
    c.mts = int(m) if m.lstrip("-").isdigit() else 0
    return c, int(vals.get("atoms", "0")), int(vals.get("virtual_sites", "0"))


def read_dir(path: str) -> System:
    """Parse a reference-format input directory into a :class:`System`."""
    ctl, natoms_ctl, nvs = read_control(os.path.join(path, "control"))
    with open(os.path.join(path, "interaction")) as f:
        lines = [ln for ln in f]
    pos = 0

    def next_tokens():
        nonlocal pos
        while pos < len(lines):
            t = _tokens(lines[pos])
            pos += 1
            if t:
                return t
        return []

    t = next_tokens()
    ntype = int(t[1])
    names, masses, labels = [None] * ntype, np.zeros(ntype), [None] * ntype
    for _ in range(ntype):
        t = next_tokens()
        k = int(t[0]) - 1
        names[k], masses[k], labels[k] = t[1], _fnum(t[2]), t[3].upper()

    def section(ntyp):
        t = next_tokens()
        n = int(t[1])
        out = []
        for _ in range(n):
            t = next_tokens()
            types = tuple(int(x) - 1 for x in t[:ntyp])
            x, v = _read_table(os.path.join(path, t[ntyp]))
            out.append(Table(types, x, v, t[ntyp]))
        return out

    bt, at, tt, nbt, ot = section(2), section(3), section(4), section(2), section(4)
    with open(os.path.join(path, "coordinate")) as f:
        cl = f.read().splitlines()
    title = (cl[0].split() or ["x"])[0]
    box = np.array([_fnum(x) for x in cl[3].split()[:3]])
    nmol = int(cl[9].split()[1])
    toks = " ".join(cl[10:]).split()
    p = 0
    itype, s, v, jbond, msz, mnm = [], [], [], [], [], []
    for _ in range(nmol):
        na = int(toks[p]); mnm.append(toks[p + 3]); p += 4
        msz.append(na)
        for _ in range(na):
            ityp, nb = int(toks[p + 1]), int(toks[p + 2])
            s.append([_fnum(x) for x in toks[p + 3:p + 6]]); p += 6
            # tokens after the 3 coordinates but before the velocities (e.g. a trailing label)
            while True:
                try:
                    _fnum(toks[p]); break
                except ValueError:
                    p += 1
            v.append([_fnum(x) for x in toks[p:p + 3]]); p += 3
            jbond.append([int(x) - 1 for x in toks[p:p + nb]]); p += nb
            itype.append(ityp - 1)
    sysm = System(box=box, type_names=names, type_mass=masses, type_label=labels,
                  itype=np.array(itype, dtype=np.int32), s=np.array(s), v=np.array(v),
                  jbond=jbond, mol_sizes=msz, mol_names=mnm, bond_tables=bt, angle_tables=at,
                  torsion_tables=tt, nb_tables=nbt, oop_tables=ot, control=ctl, title=title)
    if nvs > 0 and os.path.exists(os.path.join(path, "virtual")):
        with open(os.path.join(path, "virtual")) as f:
            vl = [x for x in f.read().splitlines()]
        nv = int(vl[2].split()[0])
        tk = " ".join(vl[5:]).split()
        p = 0
        vt, vc, va = [], [], []
        for _ in range(nv):
            _, typ, na, cen = (int(x) for x in tk[p:p + 4]); p += 4
            vt.append(typ - 1); vc.append(cen - 1)
            va.append(np.array([int(x) - 1 for x in tk[p:p + na]], dtype=np.int32)); p += na
        sysm.vs_type = np.array(vt, dtype=np.int32)
        sysm.vs_center = np.array(vc, dtype=np.int32)
        sysm.vs_atoms = va
        vb = os.path.join(path, "virtbonds")
        if os.path.exists(vb):
            sysm.virtbonds = np.loadtxt(vb, dtype=np.int64, ndmin=2)[:, :2].astype(np.int32) - 1
        va_ = os.path.join(path, "virtangles")
        if os.path.exists(va_):
            sysm.virtangles = np.loadtxt(va_, dtype=np.int64, ndmin=2)[:, :3].astype(np.int32) - 1
    return sysm


# ----------------------------------------------------------------------------------------------
# compact archive of a System (numpy .npz, no pickled objects)


def to_arrays(sysm: System, f32=False) -> dict:
    """Flat numpy arrays describing a System (inverse: :func:`from_arrays`).
    With f32, positions, velocities and tables are kept in single precision,
    which is exactly what the reference's REAL*4 readers see."""
    ft = np.float32 if f32 else np.float64
    out = {
        "box": np.asarray(sysm.box, np.float64), "type_mass": np.asarray(sysm.type_mass, np.float64),
        "type_names": np.array(sysm.type_names), "type_label": np.array(sysm.type_label),
        "itype": np.asarray(sysm.itype, np.int32), "s": np.asarray(sysm.s, ft),
        "v": np.asarray(sysm.v, ft),
        "jbond_n": np.array([len(j) for j in sysm.jbond], np.int32),
        "jbond": np.array([x for j in sysm.jbond for x in j], np.int32),
        "mol_sizes": np.asarray(sysm.mol_sizes, np.int32), "mol_names": np.array(sysm.mol_names),
    }
    c = sysm.control
    for k, val in vars(c).items():
        out["ctl_" + k] = np.array(val)
    for kind in ("bond", "angle", "torsion", "nb", "oop"):
        tabs = getattr(sysm, kind + "_tables")
        out[f"{kind}_n"] = np.array(len(tabs), np.int32)
        for q, t in enumerate(tabs):
            out[f"{kind}{q}_types"] = np.asarray(t.types, np.int32)
            out[f"{kind}{q}_x"] = np.asarray(t.x, ft)
            out[f"{kind}{q}_v"] = np.asarray(t.v, ft)
            out[f"{kind}{q}_name"] = np.array(t.name)
    if sysm.nvs:
        out["vs_type"] = np.asarray(sysm.vs_type, np.int32)
        out["vs_center"] = np.asarray(sysm.vs_center, np.int32)
        out["vs_n"] = np.array([len(a) for a in sysm.vs_atoms], np.int32)
        out["vs_atoms"] = np.concatenate(sysm.vs_atoms).astype(np.int32)
        if sysm.virtbonds is not None:
            out["virtbonds"] = np.asarray(sysm.virtbonds, np.int32)
        if sysm.virtangles is not None:
            out["virtangles"] = np.asarray(sysm.virtangles, np.int32)
    return out


def from_arrays(a) -> System:
    off = np.concatenate([[0], np.cumsum(a["jbond_n"])])
    jb = [list(map(int, a["jbond"][off[i]:off[i + 1]])) for i in range(len(a["jbond_n"]))]
    c = Control()
    for k in vars(c):
        if "ctl_" + k in a:
            val = a["ctl_" + k].item()
            setattr(c, k, type(getattr(c, k))(val))
    sysm = System(box=np.asarray(a["box"]), type_names=[str(x) for x in a["type_names"]],
                  type_mass=np.asarray(a["type_mass"]), type_label=[str(x) for x in a["type_label"]],
                  itype=np.asarray(a["itype"], np.int32), s=np.asarray(a["s"], np.float64),
                  v=np.asarray(a["v"], np.float64), jbond=jb,
                  mol_sizes=[int(x) for x in a["mol_sizes"]], mol_names=[str(x) for x in a["mol_names"]], control=c)
    for kind in ("bond", "angle", "torsion", "nb", "oop"):
        tabs = []
        for q in range(int(a[f"{kind}_n"])):
            tabs.append(Table(tuple(int(x) for x in a[f"{kind}{q}_types"]), np.asarray(a[f"{kind}{q}_x"], np.float64),
                              np.asarray(a[f"{kind}{q}_v"], np.float64), str(a[f"{kind}{q}_name"])))
        setattr(sysm, kind + "_tables", tabs)
    if "vs_type" in a:
        sysm.vs_type = np.asarray(a["vs_type"], np.int32)
        sysm.vs_center = np.asarray(a["vs_center"], np.int32)
        vo = np.concatenate([[0], np.cumsum(a["vs_n"])])
        sysm.vs_atoms = [np.asarray(a["vs_atoms"][vo[i]:vo[i + 1]], np.int32) for i in range(len(a["vs_n"]))]
        sysm.virtbonds = np.asarray(a["virtbonds"], np.int32) if "virtbonds" in a else None
        sysm.virtangles = np.asarray(a["virtangles"], np.int32) if "virtangles" in a else None
    return sysm
