"""Writes the golden fixtures in tests/golden/ from the reference's own routines.

Runs only where the reference sources are present (/root/reference): builds
oracle/_ref with oracle/Makefile.ref, then for every case writes the inputs in
the reference's file formats, runs oracle/_ref/ref_harness (the reference's
subroutines driven by this repo's harness), and stores inputs and outputs as
.npz data: forces, energies and virial after one NEW_FORCE, hashes and counts
of the neighbour pair sets, exclusion lists and enumerated terms, and states
after short runs.  The atomisticPA case also keeps, as numbers, the expected
coordinates and velocities of the first three atoms that the reference's
testsuite compares against (testsuite/tests.py:47-68).  comb_oop is a
generated comb polymer whose branch atoms carry out-of-plane terms, and
comb_oop_only the same with every other table zero (out-of-plane forces alone);
hybridPA_mts runs hybridPA through MTS_LOOP's force (MTS_FORCE, NMTS 2) for 12 steps;
cg_ortho is a melt in an orthorhombic box; trz_cg2000 keeps the trajectory
frame and restart file the reference's WRITETRJ writes after 10 steps of cg2000;
hybridPA_npt runs the testsuite's hybridPA input as shipped (NPT: SCALEBP every 10
steps), and hybridPA_npt_expected keeps the testsuite's own NPT expectations
(testsuite/tests.py:19-39) as numbers.

Inputs are stored in single precision, which is what the reference's REAL*4
readers see, so the oracle and the reference start from identical numbers.

    python tests/golden/make_golden.py [case ...]
    python tests/golden/make_golden.py cg10k_nve_drift STATE.npz   (a state dump: s nm, v reduced)
"""
from __future__ import annotations

import ast
import hashlib
import os
import re
import resource
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from ibisco_amd import system, units  # noqa: E402

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def pair_hash(p):
    p = np.ascontiguousarray(np.asarray(p, dtype="<i4").reshape(-1, 2))
    return hashlib.sha256(p.tobytes()).hexdigest()


def canonical_pairs(p):
    p = np.sort(np.asarray(p, np.int64), axis=1)
    order = np.lexsort((p[:, 1], p[:, 0]))
    return p[order].astype(np.int32)


def item_hashes(pairs, nitems):
    """Per item, an 8-byte BLAKE2b digest of its ascending partner list (both directions of every pair)."""
    p = np.asarray(pairs, np.int64).reshape(-1, 2)
    d = np.concatenate([p, p[:, ::-1]])
    d = d[np.lexsort((d[:, 1], d[:, 0]))]
    cut = np.searchsorted(d[:, 0], np.arange(nitems + 1))
    part = d[:, 1].astype("<i4")
    out = np.zeros(nitems, np.uint64)
    for i in range(nitems):
        out[i] = int.from_bytes(hashlib.blake2b(part[cut[i]:cut[i + 1]].tobytes(), digest_size=8).digest(), "little")
    return out


def list_hash(counts, flat):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(np.asarray(counts, "<i4")).tobytes())
    h.update(np.ascontiguousarray(np.asarray(flat, "<i4")).tobytes())
    return h.hexdigest()


def _stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_harness(d, *args):
    env = dict(os.environ, OMP_NUM_THREADS="8", OMP_STACKSIZE="256M")
    r = subprocess.run([HARNESS, *args], cwd=d, env=env, capture_output=True, text=True, preexec_fn=_stack)
    if r.returncode != 0:
        raise RuntimeError(f"harness {args} failed in {d}: {r.stderr[-500:]}")


def parse_force_dump(path):
    b = open(path, "rb").read()
    off = 0

    def rd(dt, n):
        nonlocal off
        a = np.frombuffer(b, dt, n, off)
        off += a.nbytes
        return a

    nitems, natoms = (int(x) for x in rd("<i4", 2))
    out = {"nitems": nitems, "natoms": natoms}
    out["forces"] = rd("<f4", 3 * nitems).reshape(-1, 3).copy()
    out["energies"] = rd("<f8", 15).copy()
    out["pt"] = rd("<f4", 6).copy()
    rd("<f4", 3 * nitems)
    for g in (0, 1):
        n = int(rd("<i4", 1)[0])
        p = rd("<i4", 2 * n).reshape(-1, 2) - 1
        cp = canonical_pairs(p)
        out[f"pairs{g}_count"] = n
        out[f"pairs{g}_hash"] = pair_hash(cp)
        out[f"pairs{g}_items"] = item_hashes(cp, nitems)
    conn = rd("<i4", nitems).copy()
    flat = rd("<i4", int(conn.sum())) - 1
    out["excl_hash"] = list_hash(conn, flat)
    n = int(rd("<i4", 1)[0])
    ang = rd("<i4", 3 * n).reshape(-1, 3) - 1
    out["angles_count"] = n
    out["angles_hash"] = pair_hash(ang.reshape(-1, 2)) if n else ""
    n = int(rd("<i4", 1)[0])
    tor = rd("<i4", 4 * n).reshape(-1, 4) - 1
    out["torsions_count"] = n
    out["torsions_hash"] = pair_hash(tor.reshape(-1, 2)) if n else ""
    return out


def parse_state(path):
    b = open(path, "rb").read()
    nitems, natoms = np.frombuffer(b, "<i4", 2)
    s = np.frombuffer(b, "<f4", 3 * nitems, 8).reshape(-1, 3)
    v = np.frombuffer(b, "<f4", 3 * natoms, 8 + 12 * nitems).reshape(-1, 3)
    return s.copy(), v.copy()


def single(sysm):
    """The system as the reference's REAL*4 readers see it."""
    return system.from_arrays(system.to_arrays(sysm, f32=True))


def testsuite_case(name):
    d = tempfile.mkdtemp(prefix="ibx_ts_")
    src = os.path.join(REF, "testsuite", name)
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), d)
        os.chmod(os.path.join(d, f), 0o644)
    with open(os.path.join(d, "control"), "a") as fh:
        fh.write("mts 0\n")   # rev 79 needs it (RDCONTROL.f90:485-497); the shipped files say "MTS N"
    sysm = system.read_dir(d)
    shutil.rmtree(d)
    return sysm


def testsuite_expected():
    """The atomistic expected arrays of testsuite/tests.py:47-68 (numbers only)."""
    txt = open(os.path.join(REF, "testsuite", "tests.py")).read()
    block = txt[txt.index("class reference_atomistic_data"):txt.index("class AtomisticTest")]
    out = {}
    for key in ("ref_coords_short", "ref_velos_short", "ref_coords_long", "ref_velos_long"):
        seg = block[block.index(key):]
        seg = seg[seg.index("[["):seg.index("]]") + 2]
        nums = [float(x) for x in re.findall(r"-?\d+\.\d+(?:[eE][-+]?\d+)?", seg)]
        out[key] = np.array(nums).reshape(3, 3)
    return out


def ortho_case():
    """A CG melt in an orthorhombic box (8.0 x 7.5 x 6.25 nm, 8 beads/nm^3): 4 x 4 x 3 reference cells."""
    return system.cg_melt(nchains=60, nbeads=50, seed=13, avoid=0.8, dt=0.002, box=(8.0, 7.5, 6.25))


def mts_case():
    """hybridPA stepped by MTS_LOOP's force (NMTS 2: explicit steps 1-3 of every 5, extrapolated 4 and 5,
    step 5 with the MOD = 0 backward extrapolation), NVE, list rebuilt every 5 steps."""
    s = testsuite_case("hybridPA")
    s.control.ensemble = "NVE"
    s.control.mts = 2
    s.control.nupdate = 5
    return s


def virt_case():
    """hybridPA with a virtbonds and a virtangles file (RDVIRTBONDS.f90:42-64, RDVIRTANGLES.f90:42-50): each
    bead's nearest virtual site within 0.6 nm (unwrapped, as bonded terms see them) excluded from it, and
    angles bead-site-bead over the two nearest beads of sites that have two within 0.6 nm, with harmonic
    angle tables (theta0 120 degrees, 40 kJ/mol/rad^2) added for each type triple they use.  NVE."""
    s = testsuite_case("hybridPA")
    s.control.ensemble = "NVE"
    lab = np.array(s.type_label)[s.itype]
    vs_types = set(int(t) for t in s.vs_type)
    beads = np.array([i for i in range(s.natoms) if lab[i] == "B" and int(s.itype[i]) not in vs_types])
    vpos = s.s[s.vs_center]
    d = np.linalg.norm(vpos[:, None, :] - s.s[beads][None, :, :], axis=2)   # unwrapped, no minimum image
    vb, va = [], []
    for k in range(s.nvs):
        near = np.argsort(d[k])[:2]
        near = [int(beads[q]) for q in near if d[k, q] < 0.6]
        if near:
            vb.append((s.natoms + k, near[0]))
        if len(near) == 2 and len(va) < 200:
            va.append((near[0], s.natoms + k, near[1]))
    s.virtbonds = np.array(vb, np.int32)
    s.virtangles = np.array(va, np.int32)
    triples = sorted({(int(s.itype[i]), int(s.vs_type[j - s.natoms]), int(s.itype[k])) for i, j, k in va})
    have = {tuple(t.types) for t in s.angle_tables}
    x, v = system.harmonic_angle_table(120.0, 40.0)
    for t in triples:
        if t not in have and t[::-1] not in have:
            s.angle_tables.append(system.Table(t, x, v, f"vang{len(s.angle_tables)}"))
            have.add(t)
    return s


def collinear_case():
    """A UA torsion melt whose chains end in a head type (1) and a tail type (2) without angle tables
    (their end angles only exclude, SETLIS.f90) but with torsion tables, and in which chain 0's first three
    atoms and chain 1's last three atoms lie exactly on a line (coordinates exact in single precision):
    the torsions there take the degenerate branches of TORSIONS.inc:47-59 (RM = 0) and :75-87 (RN = 0)."""
    s = system.ua_melt(nchains=20, nbeads=100, seed=5, avoid=0.6)
    n = s.natoms
    starts = np.cumsum([0] + list(s.mol_sizes[:-1]))
    s.itype = s.itype.copy()
    s.itype[starts] = 1
    s.itype[starts + np.array(s.mol_sizes) - 1] = 2
    s.type_names = list(s.type_names) + ["CH3h", "CH3t"]
    s.type_mass = np.concatenate([s.type_mass, [s.type_mass[0], s.type_mass[0]]])
    s.type_label = list(s.type_label) + [s.type_label[0], s.type_label[0]]
    nb0, b0, t0 = s.nb_tables[0], s.bond_tables[0], s.torsion_tables[0]
    s.nb_tables = [system.Table((a, b), nb0.x, nb0.v, f"nb{a}{b}") for a in range(3) for b in range(a, 3)]
    s.bond_tables = [system.Table((a, 0), b0.x, b0.v, f"bond{a}") for a in range(3)]
    s.torsion_tables = [system.Table((0, 0, 0, 0), t0.x, t0.v, "tors0"),
                        system.Table((1, 0, 0, 0), t0.x, t0.v, "torsh"),
                        system.Table((0, 0, 0, 2), t0.x, t0.v, "tors_t")]
    # The reference breaks the line by scaling one x coordinate by 0.99, so neither line may run along x:
    # chain 0's head along (1, 1, 0), chain 1's tail along y.
    q = 1.0 / 64.0   # exact in single precision
    a0 = int(starts[0])
    base = np.round(s.s[a0 + 1] / q) * q
    for m in range(3):
        s.s[a0 + m] = base + (m - 1) * np.array([6 * q, 6 * q, 0.0])   # 0.133 nm apart
    e1 = int(starts[1] + s.mol_sizes[1] - 1)   # chain 1's last atom
    base = np.round(s.s[e1 - 1] / q) * q
    for m in range(3):
        s.s[e1 - 2 + m] = base + (m - 1) * np.array([0.0, 10 * q, 0.0])   # 0.156 nm apart
    return s


def drift_fit(etot, dt_ps):
    """Slope (per ps) of a straight-line fit to a total-energy series sampled every step."""
    t = np.arange(1, len(etot) + 1) * dt_ps
    return float(np.polyfit(t, etot, 1)[0])


def drift_case(state_path, nsteps=1000):
    """The 10,000-bead CG melt (200 x 50, 8 beads/nm^3, LJ 0.47 nm / 3.5 kJ/mol to 1.5 nm, list 1.8 nm every 10
    steps, 5 fs, NVE) from an equilibrated state (positions nm, velocities reduced, as a state dump of the
    device path holds them), run for ``nsteps`` by the reference's own routines: its per-step energies and the
    drift of its total energy (V_NB + V_BOND + V_ANGLE + EK, reduced units) per ps."""
    from ibisco_amd import workloads
    sysm = workloads.cg_melt_10k()
    st = np.load(state_path)
    sysm.s = st["s"][:sysm.natoms].astype(np.float64)
    sysm.v = units.vel_from_reduced(st["v"].astype(np.float64))
    sysm = single(sysm)
    d = tempfile.mkdtemp(prefix="ibx_drift_")
    try:
        system.write_dir(sysm, d)
        run_harness(d, "run", str(nsteps))
        obs = np.fromfile(os.path.join(d, "ref_obs.bin"), "<f8").reshape(nsteps, -1)
        etot = obs[:, 0:15].sum(1) + obs[:, 21] + obs[:, 22]
        slope = drift_fit(etot, sysm.control.dt)
        data = {"ref_obs": obs, "ref_etot": etot, "ref_drift_per_ps": np.array(slope), "nsteps": np.array(nsteps)}
        data.update({"sys_" + k: v for k, v in system.to_arrays(sysm, f32=True).items()})
        np.savez_compressed(os.path.join(HERE, "cg10k_nve_drift.npz"), **data)
        print(f"cg10k_nve_drift: {sysm.natoms} beads, {nsteps} steps, reference drift {slope:.4e} per ps "
              f"(E from {etot[0]:.6e} to {etot[-1]:.6e})")
    finally:
        shutil.rmtree(d, ignore_errors=True)


def trz_case(nsteps=10):
    """The reference's own WRITETRJ (TRZ frame twice, restart, config.xyz) after nsteps of cg2000: its files as
    bytes and the values it wrote them from (trz_cg2000.npz)."""
    sysm = single(system.cg_melt(nchains=40, nbeads=50, seed=3, avoid=0.8))
    d = tempfile.mkdtemp(prefix="ibx_trz_")
    try:
        system.write_dir(sysm, d)
        run_harness(d, "trz", str(nsteps))
        b = open(os.path.join(d, "ref_trz_vals.bin"), "rb").read()
        natoms, ntrj = np.frombuffer(b, "<i4", 2)
        off = 8
        pt = np.frombuffer(b, "<f8", 6, off); off += 48
        epot, ek, mktemp = np.frombuffer(b, "<f8", 3, off); off += 24
        box_dt_init = np.frombuffer(b, "<f8", 5, off); off += 40
        s = np.frombuffer(b, "<f4", 3 * natoms, off).reshape(-1, 3); off += 12 * natoms
        v = np.frombuffer(b, "<f4", 3 * natoms, off).reshape(-1, 3)
        data = dict(trj=np.frombuffer(open(os.path.join(d, "s-md.trj"), "rb").read(), np.uint8),
                    restart=np.frombuffer(open(os.path.join(d, "restart"), "rb").read(), np.uint8),
                    natoms=natoms, ntrj=ntrj, nsteps=nsteps, pt=pt, epot=epot, ek=ek, mktemp=mktemp,
                    box=box_dt_init[:3], dt=box_dt_init[3], initime=box_dt_init[4], s=s.copy(), v=v.copy(),
                    title=np.array(sysm.title), mol_sizes=np.array(sysm.mol_sizes),
                    mol_names=np.array(sysm.mol_names), itype=np.array(sysm.itype, np.int32),
                    nbonds=np.array([len(j) for j in sysm.jbond], np.int32),
                    jbond=np.array([(list(j) + [-1] * 4)[:4] for j in sysm.jbond], np.int32))
        np.savez_compressed(os.path.join(HERE, "trz_cg2000.npz"), **data)
        print(f"trz_cg2000: {natoms} atoms, {len(data['trj'])} TRZ bytes, {len(data['restart'])} restart bytes")
    finally:
        shutil.rmtree(d, ignore_errors=True)


def hybrid_npt_expected():
    """The reference testsuite's own NPT expectations for hybridPA (testsuite/tests.py:19-39): coordinates and
    velocities of the first three atoms and the box after 10 and 40 steps, as MDAnalysis reads them from the
    TRZ file (angstrom, angstrom/ps)."""
    src = open(os.path.join(REF, "testsuite", "tests.py")).read()
    body = src[src.index("class reference_hybrid_data"):src.index("class reference_atomistic_data")]
    out = {}
    for key in ("ref_coords_short", "ref_velos_short", "ref_box_short", "ref_coords_long", "ref_velos_long",
                "ref_box_long"):
        m = re.search(key + r"\s*=\s*np\.array\((\[.*?\])\s*(?:,\s*dtype=np\.float32)?\)", body, re.S)
        out[key] = np.array(ast.literal_eval(m.group(1)), dtype=np.float32)
    return out


def make_case(name, sysm, runs=()):
    sysm = single(sysm)
    d = tempfile.mkdtemp(prefix="ibx_gold_")
    try:
        system.write_dir(sysm, d)
        run_harness(d, "force")
        res = parse_force_dump(os.path.join(d, "ref_dump.bin"))
        data = {"ref_" + k: np.array(v) for k, v in res.items()}
        for n in runs:
            run_harness(d, "run", str(n))
            s, v = parse_state(os.path.join(d, "ref_state.bin"))
            obs = np.fromfile(os.path.join(d, "ref_obs.bin"), "<f8").reshape(n, -1)
            keep = min(len(v), 200)
            data[f"run{n}_s"] = s[:keep]
            data[f"run{n}_v"] = v[:keep]
            data[f"run{n}_obs"] = obs
            data[f"run{n}_box"] = np.fromfile(os.path.join(d, "ref_box.bin"), "<f8")
        data.update({"sys_" + k: v for k, v in system.to_arrays(sysm, f32=True).items()})
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **data)
        print(f"{name}: {res['nitems']} items, {res['pairs0_count']} + {res['pairs1_count']} pairs")
    finally:
        shutil.rmtree(d, ignore_errors=True)


def main(only=()):
    subprocess.run(["make", "-s", "-j8", "-f", os.path.join(ROOT, "oracle", "Makefile.ref")], check=True)
    if only:
        for name in only:
            if name == "comb_oop":
                make_case("comb_oop", system.comb_melt(nchains=36, seed=1), runs=(10,))
            if name == "cg_ortho":
                make_case("cg_ortho", ortho_case(), runs=(10,))
            if name == "hybridPA_mts":
                make_case("hybridPA_mts", mts_case(), runs=(12,))
            if name == "cg10k_nve_drift":   # needs the state file after the name
                drift_case(only[only.index(name) + 1])
                return
            if name == "hybridPA_virt":
                make_case("hybridPA_virt", virt_case(), runs=(10,))
            if name == "ua_collinear":
                make_case("ua_collinear", collinear_case(), runs=(10,))
            if name == "comb_oop_only":
                make_case("comb_oop_only", system.comb_melt(nchains=36, seed=1, oop_only=True), runs=(10,))
            if name == "trz_cg2000":
                trz_case()
            if name == "hybridPA_npt":
                make_case("hybridPA_npt", testsuite_case("hybridPA"), runs=(10, 40))
                np.savez_compressed(os.path.join(HERE, "hybridPA_npt_expected.npz"), **hybrid_npt_expected())
        return
    make_case("cg2000", system.cg_melt(nchains=40, nbeads=50, seed=3, avoid=0.8), runs=(10,))
    ua = system.ua_melt(nchains=20, nbeads=100, seed=5, avoid=0.6)   # 3 cells per edge at rlist 1.3
    ua.control.non_bonded = 5
    make_case("ua2000_excl5", ua)
    st = system.cg_melt(nchains=30, nbeads=50, seed=9, avoid=0.8)
    st.s = st.s + np.random.default_rng(9).normal(0.0, 0.03, st.s.shape)   # strained bonds and angles
    make_case("cg1500_strained", st)
    tb = testsuite_case("twoatombox")
    make_case("twoatombox", tb, runs=(40,))
    make_case("hybridPA", testsuite_case("hybridPA"))
    apa = testsuite_case("atomisticPA")
    make_case("atomisticPA", apa, runs=(10, 40))
    exp = testsuite_expected()
    np.savez_compressed(os.path.join(HERE, "atomisticPA_expected.npz"), **exp)
    print("atomisticPA expected arrays:", {k: v.shape for k, v in exp.items()})
    make_case("comb_oop", system.comb_melt(nchains=36, seed=1), runs=(10,))
    make_case("comb_oop_only", system.comb_melt(nchains=36, seed=1, oop_only=True), runs=(10,))
    make_case("hybridPA_mts", mts_case(), runs=(12,))
    make_case("cg_ortho", ortho_case(), runs=(10,))
    make_case("hybridPA_virt", virt_case(), runs=(10,))
    make_case("ua_collinear", collinear_case(), runs=(10,))
    trz_case()
    make_case("hybridPA_npt", testsuite_case("hybridPA"), runs=(10, 40))
    np.savez_compressed(os.path.join(HERE, "hybridPA_npt_expected.npz"), **hybrid_npt_expected())


if __name__ == "__main__":
    main(sys.argv[1:])
