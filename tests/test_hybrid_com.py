"""Centre-of-mass virtual sites in the oracle (VIRTUAL_DEF.f90:29-40, DISTRIBUTE_VSFORCE.f90:12-41).

The reference's COM branch zeroes only the x accumulator (VIRTUAL_DEF.f90:31),
so its y and z start from whatever the previous site (or the stack) left and no
reference output of a COM system is usable as a fixture.  The oracle computes
the true centre of mass; these checks pin it by what the mathematics fixes:

* every site sits at the mass-weighted mean of its members' unwrapped
  positions (with two atom masses, so a plain mean fails);
* with the atoms' own interactions switched off, an atom feels only its site's
  force times m_j/M, so the oracle's atom forces must equal minus the
  finite-difference derivative of the oracle's energy with respect to the
  atom's position (a missing or wrong mass fraction, or a site placed
  elsewhere, fails by a large factor);
* the forces on a site's members sum to the force the site received.
"""
import numpy as np

from ibisco_amd import system


def _hybrid(**kw):
    return system.hybrid_melt(3, 8, chain_len=24, seed=2, avoid=0.8, light=True, **kw)


def test_sites_at_mass_weighted_centres(oracle_mod):
    s = _hybrid()
    o = oracle_mod.Oracle(s)
    o.new_force()
    pos, _, r = o.state()
    m = s.type_mass[s.itype]
    com = np.array([(m[a, None] * s.s[a]).sum(0) / m[a].sum() for a in s.vs_atoms])
    plain = np.array([s.s[a].mean(0) for a in s.vs_atoms])
    assert np.abs(plain - com).max() > 1e-3        # the two masses make the check meaningful
    assert np.abs(pos[s.natoms:] - com).max() <= 1e-12
    # folded copy in [-L/2, L/2] (SHIFT of the site, VIRTUAL_DEF.f90:42-46)
    L = np.asarray(s.box)
    d = r[s.natoms:] - com
    assert np.abs(d - np.round(d / L) * L).max() <= 1e-12 and np.all(np.abs(r[s.natoms:]) <= L / 2 + 1e-12)


def _site_only(s):
    """The same system with bonded and atom-atom potentials set to zero (exclusions kept)."""
    for t in s.bond_tables + s.angle_tables:
        t.v = np.zeros_like(t.v)
    for t in s.nb_tables:
        if 1 not in t.types and 2 not in t.types:
            t.v = np.zeros_like(t.v)
    return s


def test_atom_forces_are_energy_derivatives_through_sites(oracle_mod):
    s = _site_only(_hybrid(dr_nb=1e-4))
    o = oracle_mod.Oracle(s)
    o.new_force()
    f = o.forces()

    def energy(pos):
        o.set_state(pos)
        o.new_force()
        return float(o.obs()["vnb"].sum())

    h = 1e-6
    rng = np.random.default_rng(0)
    atoms = rng.choice(np.concatenate(s.vs_atoms), 24, replace=False)
    got, fd = [], []
    for a in atoms:
        for k in range(3):
            p = s.s.copy()
            p[a, k] += h
            ep = energy(p)
            p[a, k] -= 2 * h
            em = energy(p)
            fd.append(-(ep - em) / (2 * h))
            got.append(f[a, k])
    got, fd = np.array(got), np.array(fd)
    scale = np.abs(fd).max()
    assert scale > 1.0
    # piecewise-linear energy against lerped centred-difference forces: O(dr V'') apart
    assert np.sqrt(np.mean((got - fd) ** 2)) <= 1e-2 * np.sqrt(np.mean(fd ** 2))
    assert np.abs(got - fd).max() <= 2e-2 * scale


def test_member_forces_sum_to_site_force(oracle_mod):
    s = _site_only(_hybrid())
    o = oracle_mod.Oracle(s)
    o.new_force()
    f = o.forces()
    m = s.type_mass[s.itype]
    for a in s.vs_atoms[:10]:
        fa = f[a]
        # every member carries the same force per unit mass (F_vs / M)
        per_mass = fa / m[a, None]
        assert np.allclose(per_mass, per_mass[0], rtol=1e-12, atol=1e-12 * np.abs(per_mass).max())
    # total force vanishes (Newton's third law through the sites)
    assert np.abs(f[:s.natoms].sum(0)).max() <= 1e-9 * np.abs(f).max()
