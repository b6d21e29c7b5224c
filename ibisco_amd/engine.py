"""Python host binding of libibx.so (ctypes over the C ABI of include/ibx.h).

There is no CPU fallback: constructing an :class:`Engine` without the built
library or without a gfx950 device raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .arrays import flatten

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libibx.so")

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class IbxTables(C.Structure):
    _fields_ = [("ntab", C.c_int), ("types", _ip), ("off", _ip), ("nrow", _ip), ("x", _dp), ("v", _dp)]


class IbxSystemDesc(C.Structure):
    _fields_ = [
        ("natoms", C.c_int), ("nvs", C.c_int), ("ntype", C.c_int),
        ("mass", _dp), ("label", _ip), ("itype", _ip), ("nbonds", _ip), ("jbond", _ip),
        ("s", _dp), ("v", _dp), ("box", C.c_double * 3),
        ("bond", IbxTables), ("angle", IbxTables), ("tors", IbxTables), ("nb", IbxTables), ("oop", IbxTables),
        ("vs_type", _ip), ("vs_center", _ip), ("vs_off", _ip), ("vs_atoms", _ip),
        ("nvirtbonds", C.c_int), ("virtbonds", _ip), ("nvirtangles", C.c_int), ("virtangles", _ip),
        ("ensemble", C.c_int), ("dt", C.c_double), ("temp_in", C.c_double), ("fac", C.c_double),
        ("rcut_atom", C.c_double), ("rlist_atom", C.c_double), ("rcut_bead", C.c_double),
        ("rlist_bead", C.c_double), ("nonbexc_atom", C.c_int), ("nonbexc_bead", C.c_int),
        ("nupdate", C.c_int), ("halt_drift", C.c_int), ("nmts", C.c_int), ("mts_quirk", C.c_int),
        ("pressure", C.c_double), ("pfac", C.c_double),
    ]


class IbxServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_cohorts", C.c_int), ("max_cohort_jobs", C.c_int),
                ("chunk_steps", C.c_int), ("rdf_nbin", C.c_int), ("rdf_dr", C.c_double), ("step_band", C.c_double),
                ("spare_places", C.c_int), ("class_ratio", C.c_double)]


class IbxJobResult(C.Structure):
    _fields_ = [("status", C.c_int), ("natoms", C.c_int), ("nsteps", C.c_int), ("cohort", C.c_int),
                ("cohort_jobs", C.c_int), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("mean_pressure", C.c_double), ("obs", C.c_double * 32),
                ("sums", C.c_double * 32), ("error", C.c_char * 256)]


EXPORTS = ("ibx_create", "ibx_load_dir", "ibx_destroy", "ibx_last_error", "ibx_step", "ibx_step_graph",
           "ibx_compute_forces", "ibx_rebuild_lists", "ibx_get_observables", "ibx_get_averages",
           "ibx_get_forces", "ibx_get_state", "ibx_set_state", "ibx_set_nb_table", "ibx_set_thermostat",
           "ibx_get_pairs", "ibx_rdf_enable", "ibx_rdf_get", "ibx_get_timing", "ibx_set_timing",
           "ibx_get_info", "ibx_retire", "ibx_server_create", "ibx_server_submit", "ibx_server_wait",
           "ibx_server_stats", "ibx_server_destroy", "ibx_server_last_error", "ibx_estimate_bytes",
           "ibx_set_skin_tuning", "ibx_get_skin", "ibx_set_respa", "ibx_get_counters", "ibx_step_noalloc",
           "ibx_assign", "ibx_get_capacity", "ibx_server_timing", "ibx_set_last_step", "ibx_trz_write_frame",
           "ibx_restart_write", "ibx_write_trz", "ibx_write_restart", "ibx_output_last_error", "ibx_get_box",
           "ibx_set_tiles", "ibx_get_tiles")

_lib = None


def load_library(path=LIB_PATH):
    """Loads libibx.so and declares its signatures (no device work)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m ibisco_amd.build`")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.ibx_create.argtypes = [C.c_int, C.POINTER(C.POINTER(IbxSystemDesc)), C.POINTER(vp)]
    L.ibx_load_dir.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.ibx_destroy.argtypes = [vp]
    L.ibx_destroy.restype = None
    L.ibx_last_error.argtypes = [vp]
    L.ibx_last_error.restype = C.c_char_p
    for nm in ("ibx_step", "ibx_step_graph", "ibx_step_noalloc"):
        getattr(L, nm).argtypes = [vp, C.c_int, C.c_int]
    L.ibx_compute_forces.argtypes = [vp, C.c_int]
    L.ibx_rebuild_lists.argtypes = [vp]
    L.ibx_get_observables.argtypes = [vp, C.c_int, _dp]
    L.ibx_get_box.argtypes = [vp, C.c_int, _dp]
    L.ibx_get_averages.argtypes = [vp, C.c_int, _dp, C.c_int]
    L.ibx_get_forces.argtypes = [vp, C.c_int, _dp]
    L.ibx_get_state.argtypes = [vp, C.c_int, _dp, _dp]
    L.ibx_set_state.argtypes = [vp, C.c_int, _dp, _dp]
    L.ibx_set_nb_table.argtypes = [vp, C.c_int, C.c_int, C.c_int, _dp, _dp]
    L.ibx_set_thermostat.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double]
    L.ibx_get_pairs.argtypes = [vp, C.c_int, C.c_int, _ip, C.c_int64, C.POINTER(C.c_int64)]
    L.ibx_rdf_enable.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    L.ibx_rdf_get.argtypes = [vp, C.c_int, _dp, C.c_int]
    L.ibx_get_timing.argtypes = [vp, _dp]
    L.ibx_set_timing.argtypes = [vp, C.c_int]
    L.ibx_get_info.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
    L.ibx_get_counters.argtypes = [vp, C.POINTER(C.c_int64)]
    L.ibx_retire.argtypes = [vp, C.c_int]
    L.ibx_assign.argtypes = [vp, C.c_int, C.POINTER(IbxSystemDesc), C.c_int]
    L.ibx_set_last_step.argtypes = [vp, C.c_int, C.c_int]
    L.ibx_trz_write_frame.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp,
                                      C.c_double, C.c_double, C.c_double, _dp, _dp]
    L.ibx_restart_write.argtypes = [C.c_char_p, C.c_char_p, C.c_double, _dp, C.c_int, _ip, _ip, _ip, _dp, _dp,
                                    C.c_int, _ip, C.POINTER(C.c_char_p)]
    L.ibx_write_trz.argtypes = [vp, C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double]
    L.ibx_write_restart.argtypes = [vp, C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.c_double, C.c_int, _ip,
                                    C.POINTER(C.c_char_p)]
    L.ibx_output_last_error.argtypes = []
    L.ibx_output_last_error.restype = C.c_char_p
    L.ibx_get_capacity.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
    L.ibx_set_skin_tuning.argtypes = [vp, C.c_int]
    L.ibx_set_tiles.argtypes = [vp, _ip]
    L.ibx_get_tiles.argtypes = [vp, _ip, C.POINTER(C.c_float)]
    L.ibx_set_respa.argtypes = [vp, C.c_int]
    L.ibx_get_skin.argtypes = [vp, C.c_int, _dp]
    L.ibx_server_create.argtypes = [C.POINTER(IbxServerConfig), C.POINTER(vp)]
    L.ibx_server_submit.argtypes = [vp, C.POINTER(IbxSystemDesc), C.c_int, C.c_int, C.POINTER(C.c_int64)]
    L.ibx_server_wait.argtypes = [vp, C.c_int64, C.c_double, C.POINTER(IbxJobResult), _dp, _dp, _dp]
    L.ibx_server_stats.argtypes = [vp, _dp]
    L.ibx_server_timing.argtypes = [vp, _dp]
    L.ibx_server_destroy.argtypes = [vp]
    L.ibx_server_destroy.restype = None
    L.ibx_estimate_bytes.argtypes = [C.POINTER(IbxSystemDesc), C.c_int, C.POINTER(C.c_int64)]
    L.ibx_server_last_error.argtypes = []
    L.ibx_server_last_error.restype = C.c_char_p
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class IbxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ibx error {code}: {msg}")
        self.code = code


def make_desc(sysm, keep, mts_quirk=False):
    """ibx_system_desc for one System; arrays it points to are appended to ``keep``."""
    a = flatten(sysm)
    d = IbxSystemDesc()
    d.natoms, d.nvs, d.ntype = a["natoms"], a["nvs"], a["ntype"]

    def arr(x, t):
        x = np.ascontiguousarray(x)
        keep.append(x)
        return _p(x, t)

    d.mass, d.label = arr(a["mass"], C.c_double), arr(a["label"], C.c_int)
    d.itype, d.nbonds, d.jbond = arr(a["itype"], C.c_int), arr(a["nbonds"], C.c_int), arr(a["jbond"], C.c_int)
    d.s, d.v = arr(a["s"], C.c_double), arr(a["v"], C.c_double)
    for k in range(3):
        d.box[k] = a["box"][k]
    for kind in ("bond", "angle", "tors", "nb", "oop"):
        t = a["tables"][kind]
        setattr(d, kind, IbxTables(t["ntab"], arr(t["types"], C.c_int), arr(t["off"], C.c_int),
                                   arr(t["nrow"], C.c_int), arr(t["x"], C.c_double), arr(t["v"], C.c_double)))
    d.vs_type, d.vs_center = arr(a["vs_type"], C.c_int), arr(a["vs_center"], C.c_int)
    d.vs_off, d.vs_atoms = arr(a["vs_off"], C.c_int), arr(a["vs_atoms"], C.c_int)
    d.nvirtbonds, d.virtbonds = a["nvirtbonds"], arr(a["virtbonds"], C.c_int)
    d.nvirtangles, d.virtangles = a["nvirtangles"], arr(a["virtangles"], C.c_int)
    d.ensemble, d.dt, d.temp_in, d.fac = a["ensemble"], a["dt"], a["temp_in"], a["fac"]
    d.rcut_atom, d.rlist_atom = a["rcut_atom"], a["rlist_atom"]
    d.rcut_bead, d.rlist_bead = a["rcut_bead"], a["rlist_bead"]
    d.nonbexc_atom, d.nonbexc_bead = a["nonbexc_atom"], a["nonbexc_bead"]
    d.nupdate, d.halt_drift, d.nmts, d.mts_quirk = a["nupdate"], a["halt_drift"], a["nmts"], int(mts_quirk)
    d.pressure, d.pfac = a["pressure"], a["pfac"]
    keep.append(d)
    return d, a


OBS_SLICES = dict(vnb=(0, 3), vbond=(3, 6), vangle=(6, 9), vtors=(9, 12), voop=(12, 15), pt=(15, 21), ek=(21, 23))


def obs_dict(o):
    d = {k: o[a:b].copy() for k, (a, b) in OBS_SLICES.items()}
    d.update(temp=o[23], temp_atom=o[24], temp_bead=o[25], volume=o[26], overext=int(o[27]), nsamples=int(o[28]))
    return d


class Engine:
    """One context: a single system or a batch of independent replicas on the GPU."""

    def __init__(self, systems, mts_quirk=False, load_dir=None):
        self.L = load_library()
        self.h = C.c_void_p()
        self._keep = []
        self.arrays = []
        if load_dir is not None:
            rc = self.L.ibx_load_dir(load_dir.encode(), C.byref(self.h))
            self.nsys = 1
            self.natoms = None
        else:
            if not isinstance(systems, (list, tuple)):
                systems = [systems]
            descs = [make_desc(s, self._keep, mts_quirk) for s in systems]
            self.arrays = [a for _, a in descs]
            ptrs = (C.POINTER(IbxSystemDesc) * len(descs))(*[C.pointer(d) for d, _ in descs])
            rc = self.L.ibx_create(len(descs), ptrs, C.byref(self.h))
            self.nsys = len(descs)
            self.natoms = [a["natoms"] for a in self.arrays]
            self.nitems = [a["natoms"] + a["nvs"] for a in self.arrays]
        if rc != 0:
            raise IbxError(rc, self.L.ibx_last_error(None).decode())
        if self.natoms is None:
            info = self.info(0)
            self.natoms = [int(info[0])]
            self.nitems = [int(info[0] + info[1])]
        self.step_count = 0

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.ibx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise IbxError(rc, self.L.ibx_last_error(self.h).decode())

    def step(self, nsteps, graph=False):
        fn = self.L.ibx_step_graph if graph else self.L.ibx_step
        self._chk(fn(self.h, nsteps, self.step_count))
        self.step_count += nsteps

    def step_noalloc(self, nsteps):
        """ibx_step_noalloc: the steps of ibx_step with the library's static scratch array for the reduction
        partials (no device allocation in the call)."""
        self._chk(self.L.ibx_step_noalloc(self.h, nsteps, self.step_count))
        self.step_count += nsteps

    def compute_forces(self, rebuild=True):
        self._chk(self.L.ibx_compute_forces(self.h, int(rebuild)))

    def rebuild_lists(self):
        self._chk(self.L.ibx_rebuild_lists(self.h))

    def obs(self, sys=0):
        o = np.zeros(32)
        self._chk(self.L.ibx_get_observables(self.h, sys, _p(o, C.c_double)))
        return obs_dict(o)

    def box(self, sys=0):
        """The current box edges (nm); an NPT system's box follows its barostat."""
        b = np.zeros(3)
        self._chk(self.L.ibx_get_box(self.h, sys, _p(b, C.c_double)))
        return b

    def averages(self, sys=0, reset=False):
        o = np.zeros(32)
        self._chk(self.L.ibx_get_averages(self.h, sys, _p(o, C.c_double), int(reset)))
        return o

    def forces(self, sys=0):
        f = np.zeros((self.nitems[sys], 3))
        self._chk(self.L.ibx_get_forces(self.h, sys, _p(f, C.c_double)))
        return f

    def state(self, sys=0):
        s = np.zeros((self.nitems[sys], 3))
        v = np.zeros((self.natoms[sys], 3))
        self._chk(self.L.ibx_get_state(self.h, sys, _p(s, C.c_double), _p(v, C.c_double)))
        return s, v

    def set_state(self, s=None, v=None, sys=0):
        s = None if s is None else np.ascontiguousarray(s, np.float64)
        v = None if v is None else np.ascontiguousarray(v, np.float64)
        self._chk(self.L.ibx_set_state(self.h, sys, None if s is None else _p(s, C.c_double),
                                       None if v is None else _p(v, C.c_double)))

    def set_nb_table(self, x, v_reduced, sys=0, t=0):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v_reduced, np.float64)
        self._chk(self.L.ibx_set_nb_table(self.h, sys, t, len(x), _p(x, C.c_double), _p(v, C.c_double)))

    def set_thermostat(self, ensemble, temp_in, fac, dt):
        self._chk(self.L.ibx_set_thermostat(self.h, ensemble, temp_in, fac, dt))

    def pairs(self, sys=0, group=0):
        n = C.c_int64()
        self._chk(self.L.ibx_get_pairs(self.h, sys, group, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 2), np.int32)
        self._chk(self.L.ibx_get_pairs(self.h, sys, group, _p(out, C.c_int), n.value, C.byref(n)))
        return out[:n.value]

    def pair_count(self, sys=0, group=0):
        """Number of unordered pairs in the current list (no copy of the list itself)."""
        n = C.c_int64()
        self._chk(self.L.ibx_get_pairs(self.h, sys, group, None, 0, C.byref(n)))
        return n.value

    def rdf_enable(self, nbin, dr, every=1):
        self._chk(self.L.ibx_rdf_enable(self.h, nbin, dr, every))
        self._rdf_nbin = nbin

    def rdf(self, sys=0, reset=False):
        g = np.zeros(self._rdf_nbin)
        self._chk(self.L.ibx_rdf_get(self.h, sys, _p(g, C.c_double), int(reset)))
        return g

    def set_timing(self, on=True):
        self._chk(self.L.ibx_set_timing(self.h, int(on)))

    def timing(self):
        t = np.zeros(8)
        self._chk(self.L.ibx_get_timing(self.h, _p(t, C.c_double)))
        return dict(nonbonded=t[0], bonded=t[1], integrate=t[2], rebuild=t[3], step=t[4], nsteps_timed=t[6],
                    rebuilds=t[7])

    def info(self, sys=0):
        i = (C.c_int64 * 8)()
        self._chk(self.L.ibx_get_info(self.h, sys, i))
        return list(i)

    def counters(self):
        """Step counters (ibx_get_counters): eager steps, graph-replayed steps, graph launches, captures,
        the block period and the long block's steps."""
        i = (C.c_int64 * 8)()
        self._chk(self.L.ibx_get_counters(self.h, i))
        return dict(eager_steps=i[0], graph_steps=i[1], graph_launches=i[2], captures=i[3], period=i[4],
                    long_block=i[5])

    def retire(self, sys):
        self._chk(self.L.ibx_retire(self.h, sys))

    def assign(self, sys, sysm, step0=None):
        """Place ``sys`` (retired) taken over by system ``sysm`` from step ``step0`` (default: the context's
        current step) on (ibx_assign); its own steps count from there."""
        step0 = self.step_count if step0 is None else int(step0)
        d, a = make_desc(sysm, self._keep)
        self._chk(self.L.ibx_assign(self.h, int(sys), C.byref(d), step0))
        self.arrays[sys] = a
        self.natoms[sys] = a["natoms"]
        self.nitems[sys] = a["natoms"] + a["nvs"]

    def write_trz(self, path, frame, step, sys=0, title="ibx", ntrj=1, initime_ps=0.0):
        """Appends a TRZ frame of system ``sys`` (ibx_write_trz; frame 1 creates the file)."""
        rc = self.L.ibx_write_trz(self.h, int(sys), path.encode(), title.encode(), int(frame), int(ntrj), int(step),
                                  float(initime_ps))
        if rc != 0:
            raise IbxError(rc, self.L.ibx_output_last_error().decode())

    def write_restart(self, path, step, mol_sizes, mol_names, sys=0, title="ibx", initime_ps=0.0):
        """The restart file of system ``sys`` (ibx_write_restart)."""
        natm = np.ascontiguousarray(mol_sizes, np.int32)
        names = (C.c_char_p * len(mol_names))(*[str(n).encode() for n in mol_names])
        rc = self.L.ibx_write_restart(self.h, int(sys), path.encode(), title.encode(), int(step), float(initime_ps),
                                      len(natm), _p(natm, C.c_int), names)
        if rc != 0:
            raise IbxError(rc, self.L.ibx_output_last_error().decode())

    def set_last_step(self, sys, nsteps):
        """System ``sys`` stops by itself after its own step ``nsteps`` (< 0: no limit; ibx_set_last_step)."""
        self._chk(self.L.ibx_set_last_step(self.h, int(sys), int(nsteps)))

    def capacity(self, sys=0):
        """What place ``sys`` holds (ibx_get_capacity)."""
        i = (C.c_int64 * 16)()
        self._chk(self.L.ibx_get_capacity(self.h, sys, i))
        keys = ("items", "terms", "term_slots", "excl", "members0", "members1", "cells0", "cells1", "nb_rows",
                "nb_tabs", "nb_map", "bd_tabs", "bd_rows", "list_cap", "period")
        return {k: int(i[q]) for q, k in enumerate(keys)}

    def set_respa(self, k):
        """Nonbonded forces every k steps with weight k, bonded every step (ibx_set_respa)."""
        self._chk(self.L.ibx_set_respa(self.h, int(k)))

    def set_skin_tuning(self, on=True):
        self._chk(self.L.ibx_set_skin_tuning(self.h, int(on)))

    def set_tiles(self, layout=-1, chunk=-1, per_cu=-1):
        """Force launch layout (ibx_set_tiles): 0 one workgroup per block, 1 persistent work queues taking
        ``chunk`` blocks with ``per_cu`` workgroups per CU; -1 leaves the choice to the tile autotuner."""
        t = np.array([layout, chunk, per_cu], dtype=np.int32)
        self._chk(self.L.ibx_set_tiles(self.h, _p(t, C.c_int)))

    def tiles(self):
        """(layout in use [layout, chunk, workgroups per CU], the autotuner's ms per candidate)."""
        t = np.zeros(3, dtype=np.int32)
        ms = np.zeros(4, dtype=np.float32)
        self._chk(self.L.ibx_get_tiles(self.h, _p(t, C.c_int), _p(ms, C.c_float)))
        return [int(x) for x in t], [float(x) for x in ms]

    def skin(self, sys=0):
        d = np.zeros(2)
        self._chk(self.L.ibx_get_skin(self.h, sys, _p(d, C.c_double)))
        return d


IBX_E_PENDING = -7


def estimate_bytes(sysm, rdf_nbin=0):
    """Device bytes a one-system context of ``sysm`` allocates (ibx_estimate_bytes; no device work)."""
    L = load_library()
    keep = []
    d, _ = make_desc(sysm, keep)
    b = C.c_int64()
    rc = L.ibx_estimate_bytes(C.byref(d), int(rdf_nbin), C.byref(b))
    if rc != 0:
        raise IbxError(rc, L.ibx_server_last_error().decode())
    return b.value


class Server:
    """The serving loop of libibx (ibx_server_*): jobs are admitted by device
    footprint into batched contexts that run concurrently; see include/ibx.h."""

    def __init__(self, hbm_budget=200e9, max_cohorts=3, max_cohort_jobs=256, chunk_steps=200, rdf_nbin=150,
                 rdf_dr=0.01, step_band=1.5, spare_places=0, class_ratio=0.0):
        self.L = load_library()
        self.h = C.c_void_p()
        cfg = IbxServerConfig(int(hbm_budget), max_cohorts, max_cohort_jobs, chunk_steps, rdf_nbin, rdf_dr,
                              step_band, int(spare_places), float(class_ratio))
        rc = self.L.ibx_server_create(C.byref(cfg), C.byref(self.h))
        if rc != 0:
            raise IbxError(rc, self.L.ibx_server_last_error().decode())
        self.rdf_nbin = rdf_nbin
        self.shape = {}

    def submit(self, sysm, nsteps, keep_state=False):
        keep = []
        d, a = make_desc(sysm, keep)
        job = C.c_int64()
        rc = self.L.ibx_server_submit(self.h, C.byref(d), int(nsteps), int(keep_state), C.byref(job))
        if rc != 0:
            raise IbxError(rc, self.L.ibx_server_last_error().decode())
        self.shape[job.value] = (a["natoms"] + a["nvs"], a["natoms"], bool(keep_state))
        return job.value

    def wait(self, job, timeout=None):
        """The job's result (dict), or None if it is still running after ``timeout`` seconds."""
        nitems, natoms, ks = self.shape[job]
        res = IbxJobResult()
        g = np.zeros(max(1, self.rdf_nbin))
        s = np.zeros((nitems, 3)) if ks else None
        v = np.zeros((natoms, 3)) if ks else None
        rc = self.L.ibx_server_wait(self.h, job, -1.0 if timeout is None else float(timeout), C.byref(res),
                                    _p(g, C.c_double), None if s is None else _p(s, C.c_double),
                                    None if v is None else _p(v, C.c_double))
        if rc == IBX_E_PENDING:
            return None
        del self.shape[job]
        if rc != 0:
            raise IbxError(rc, res.error.decode() or self.L.ibx_server_last_error().decode())
        return dict(job=job, natoms=res.natoms, nsteps=res.nsteps, cohort=res.cohort, cohort_jobs=res.cohort_jobs,
                    t_submit=res.t_submit, t_admit=res.t_admit, t_done=res.t_done,
                    mean_pressure=res.mean_pressure, obs=np.array(res.obs[:]), sums=np.array(res.sums[:]),
                    rdf=g[:self.rdf_nbin].copy(), s=s, v=v)

    def stats(self):
        st = np.zeros(16)
        rc = self.L.ibx_server_stats(self.h, _p(st, C.c_double))
        if rc != 0:
            raise IbxError(rc, self.L.ibx_server_last_error().decode())
        keys = ("submitted", "completed", "failed", "queued", "running", "cohorts", "peak_running", "bytes_in_use",
                "peak_bytes", "atom_steps", "clock_s", "largest_cohort", "mean_cohort_jobs", "budget",
                "admitted_running", "atom_steps_run")
        return {k: float(st[i]) for i, k in enumerate(keys)}

    def timing(self):
        """Host seconds of the batches' threads (ibx_server_timing)."""
        t = np.zeros(8)
        rc = self.L.ibx_server_timing(self.h, _p(t, C.c_double))
        if rc != 0:
            raise IbxError(rc, self.L.ibx_server_last_error().decode())
        return dict(step_s=t[0], assign_s=t[1], collect_s=t[2], create_s=t[3], step_calls=int(t[4]),
                    assigns=int(t[5]), steps=int(t[6]))

    def close(self):
        """Runs every submitted job to its end, then frees the server."""
        if getattr(self, "h", None) and self.h.value:
            self.L.ibx_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def trz_write_frame(path, title, frame, ntrj, time_ps, box, pt, epot, ek, mktemp, s, v):
    """One TRZ frame from host arrays (ibx_trz_write_frame; reduced units, s nm and v reduced, natoms x 3)."""
    L = load_library()
    s = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
    box = np.ascontiguousarray(box, np.float64)
    pt = np.ascontiguousarray(pt, np.float64)
    rc = L.ibx_trz_write_frame(path.encode(), title.encode(), int(frame), int(ntrj), len(s), float(time_ps),
                               _p(box, C.c_double), _p(pt, C.c_double), float(epot), float(ek), float(mktemp),
                               _p(s, C.c_double), _p(v, C.c_double))
    if rc != 0:
        raise IbxError(rc, L.ibx_output_last_error().decode())


def restart_write(path, title, time_ps, box, itype, nbonds, jbond, s, v, mol_sizes, mol_names):
    """The restart file from host arrays (ibx_restart_write; itype, jbond 0-based, jbond natoms x 4)."""
    L = load_library()
    s = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
    box = np.ascontiguousarray(box, np.float64)
    itype = np.ascontiguousarray(itype, np.int32)
    nbonds = np.ascontiguousarray(nbonds, np.int32)
    jbond = np.ascontiguousarray(jbond, np.int32).reshape(-1, 4)
    natm = np.ascontiguousarray(mol_sizes, np.int32)
    names = (C.c_char_p * len(mol_names))(*[str(n).encode() for n in mol_names])
    rc = L.ibx_restart_write(path.encode(), title.encode(), float(time_ps), _p(box, C.c_double), len(s),
                             _p(itype, C.c_int), _p(nbonds, C.c_int), _p(jbond, C.c_int), _p(s, C.c_double),
                             _p(v, C.c_double), len(natm), _p(natm, C.c_int), names)
    if rc != 0:
        raise IbxError(rc, L.ibx_output_last_error().decode())
