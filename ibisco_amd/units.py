"""IBIsCO reduced units.

Restates ``UNIT.f90:9-23`` (and the constants at ``module_var.F90:14-17``) in FP64.
Internal units: length nm, energy k_B*300 K per particle, mass 12 g/mol per
particle, so time = 2.1934 ps.  Input conversions follow the readers:
velocities in nm/ps (``RDCOOR.f90:125``), time step in ps (``RDCONTROL.f90:162``),
table energies in kJ/mol (``RDINTERACT.f90:217,526``), masses in kg/mol
(``RDINTERACT.f90:110``), temperatures in K (``RDCONTROL.f90:88``).
"""
import math

KB = 1.380658e-23
NA = 6.0221367e23
PI = 3.14159265358
MMOLY = 0.012

MASSSCALE = MMOLY / NA
RSCALE = 1.0e-9
ESCALE = KB * 300.0
TEMPSCALE = ESCALE / KB
PSCALE = ESCALE / RSCALE ** 3 / 1000.0          # reduced pressure -> kPa
TIMESCALE = RSCALE * math.sqrt(MASSSCALE / ESCALE)  # reduced time -> s
VSCALE = RSCALE / TIMESCALE                     # reduced velocity -> m/s
CONV = NA * ESCALE / 1000.0                     # reduced energy -> kJ/mol
R2D = 180.0 / PI
D2R = PI / 180.0


def kjmol_to_reduced(e):
    return e * 1000.0 / NA / ESCALE


def reduced_to_kjmol(e):
    return e * NA * ESCALE / 1000.0


def mass_to_reduced(m_kgmol):
    return m_kgmol / NA / MASSSCALE


def vel_to_reduced(v_nmps):
    return v_nmps * 1.0e3 / VSCALE


def vel_from_reduced(v):
    return v * VSCALE / 1.0e3


def dt_to_reduced(dt_ps):
    return dt_ps * 1.0e-12 / TIMESCALE


def temp_to_reduced(t_k):
    return t_k / TEMPSCALE


def pressure_to_bar(p):
    """Reduced pressure (energy unit per nm^3) in bar (PSCALE is in kPa)."""
    return p * PSCALE / 100.0
