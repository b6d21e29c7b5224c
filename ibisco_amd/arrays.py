"""Flat, 0-based, reduced-unit arrays of a :class:`~ibisco_amd.system.System`.

This is the form both C ABIs take (``include/ibx.h`` and the oracle): per-type
masses and A/B labels, per-atom type/bond adjacency/positions/velocities, raw
table rows (abscissa as in the file, energy converted from kJ/mol like
``RDINTERACT.f90:217,526``), virtual-site CSR and the control scalars
(``RDCONTROL.f90`` conversions).
"""
from __future__ import annotations

import numpy as np

from . import units
from .system import MAXBONDS, System

TABLE_KINDS = ("bond", "angle", "tors", "nb", "oop")
TABLE_ARITY = {"bond": 2, "angle": 3, "tors": 4, "nb": 2, "oop": 4}


def _tables(tabs, k):
    ntab = len(tabs)
    types = np.array([t.types[:k] for t in tabs], dtype=np.int32).reshape(-1)
    nrow = np.array([len(t.x) for t in tabs], dtype=np.int32)
    off = np.zeros(ntab, dtype=np.int32)
    if ntab:
        off[1:] = np.cumsum(nrow)[:-1]
    x = np.concatenate([np.asarray(t.x, np.float64) for t in tabs]) if ntab else np.zeros(1)
    v = (np.concatenate([units.kjmol_to_reduced(np.asarray(t.v, np.float64)) for t in tabs])
         if ntab else np.zeros(1))
    return dict(ntab=ntab, types=types if ntab else np.zeros(1, np.int32), off=off if ntab else np.zeros(1, np.int32),
                nrow=nrow if ntab else np.zeros(1, np.int32), x=x, v=v)


def flatten(sysm: System):
    n = sysm.natoms
    nb = np.zeros(n, dtype=np.int32)
    jb = np.full((n, MAXBONDS), -1, dtype=np.int32)
    for i, js in enumerate(sysm.jbond):
        if len(js) > MAXBONDS:
            raise ValueError(f"atom {i + 1} has more than {MAXBONDS} bonds")
        nb[i] = len(js)
        jb[i, :len(js)] = js
    c = sysm.control
    dt = units.dt_to_reduced(c.dt)
    taut = units.dt_to_reduced(c.taut)
    out = dict(
        natoms=n, nvs=sysm.nvs, ntype=len(sysm.type_names),
        mass=np.array([units.mass_to_reduced(m) for m in sysm.type_mass], dtype=np.float64),
        label=np.array([1 if lab.upper() == "A" else 2 for lab in sysm.type_label], dtype=np.int32),
        itype=np.ascontiguousarray(sysm.itype, dtype=np.int32),
        nbonds=nb, jbond=np.ascontiguousarray(jb.reshape(-1)),
        s=np.ascontiguousarray(sysm.s, dtype=np.float64).reshape(-1),
        v=np.ascontiguousarray(units.vel_to_reduced(np.asarray(sysm.v, np.float64))).reshape(-1),
        box=np.asarray(sysm.box, dtype=np.float64),
        tables={k: _tables(getattr(sysm, f"{k if k != 'tors' else 'torsion'}_tables"), TABLE_ARITY[k])
                for k in TABLE_KINDS},
        ensemble={"NVE": 0, "NVT": 1, "NPT": 2}[c.ensemble.upper()],
        dt=dt, temp_in=units.temp_to_reduced(c.temperature), fac=dt / taut,
        # SCALEBP.f90: PRESSURE = PRESSURE0/PSCALE, PFAC = BETA*DT*LIMAVP/TAUP (RDCONTROL.f90:107-108,317-337,
        # ALLOCATEVAR.f90:17, LIMAVP = 10)
        pressure=c.pressure / units.PSCALE,
        pfac=(c.beta * units.PSCALE) * dt * 10 / units.dt_to_reduced(c.taup),
        rcut_atom=c.cutoff, rlist_atom=c.rlist, rcut_bead=c.bead_cutoff, rlist_bead=c.bead_rlist,
        nonbexc_atom=c.non_bonded, nonbexc_bead=c.non_bonded_bead,
        nupdate=c.nupdate, halt_drift=c.halt_drift, nmts=c.mts,
    )
    if sysm.nvs:
        off = np.zeros(sysm.nvs + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(a) for a in sysm.vs_atoms])
        out.update(vs_type=np.asarray(sysm.vs_type, np.int32), vs_center=np.asarray(sysm.vs_center, np.int32),
                   vs_off=off, vs_atoms=np.concatenate(sysm.vs_atoms).astype(np.int32))
    else:
        out.update(vs_type=np.zeros(1, np.int32), vs_center=np.zeros(1, np.int32),
                   vs_off=np.zeros(2, np.int32), vs_atoms=np.zeros(1, np.int32))
    vb = sysm.virtbonds if sysm.virtbonds is not None else np.zeros((0, 2), np.int32)
    va = sysm.virtangles if sysm.virtangles is not None else np.zeros((0, 3), np.int32)
    out.update(nvirtbonds=len(vb), virtbonds=np.ascontiguousarray(vb, np.int32).reshape(-1) if len(vb) else np.zeros(2, np.int32),
               nvirtangles=len(va), virtangles=np.ascontiguousarray(va, np.int32).reshape(-1) if len(va) else np.zeros(3, np.int32))
    return out
