"""Python handle on the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: this package is the checker.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import it;
the product (omegagene_amd + libog_engine.so) never does.

Each wrapper names the C function it calls; the C sources cite the reference
file:line they restate.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

CELLPAIR_DTYPE = np.dtype([("cell_id1", "<i4"), ("cell_id2", "<i4"), ("image", "<i4"), ("mask0", "<i4"),
                           ("mask1", "<i4")])


class Grid(C.Structure):
    _fields_ = [("L", C.c_float * 3), ("lb", C.c_float * 3), ("cutoff_pair", C.c_float),
                ("cutoff_pair_half", C.c_float), ("cutoff_pair2", C.c_float), ("n_cells_xyz", C.c_int * 3),
                ("L_cell_xyz", C.c_float * 3), ("n_neighbors_xyz", C.c_int * 3), ("n_columns", C.c_int),
                ("n_uni_z", C.c_int), ("L_z_uni", C.c_float), ("max_n_cells", C.c_int),
                ("max_n_cell_pairs", C.c_int), ("max_n_atoms_column", C.c_int)]


class Cells(C.Structure):
    _fields_ = [("n_cells", C.c_int), ("n_atom_array", C.c_int), ("atomids", C.POINTER(C.c_int)),
                ("crd", C.POINTER(C.c_double)), ("cell_crd", C.POINTER(C.c_int)),
                ("idx_xy_head_cell", C.POINTER(C.c_int)), ("idx_xy_head_atom", C.POINTER(C.c_int)),
                ("n_atoms_col", C.POINTER(C.c_int)), ("n_cells_z", C.POINTER(C.c_int)),
                ("column_of_atom", C.POINTER(C.c_int)), ("cap_atom_array", C.c_int), ("cap_cells", C.c_int)]


class ZD(C.Structure):
    """orc_zd_t: zero-multipole constants of the selected order."""
    _fields_ = [("cutoff", C.c_float), ("alpha", C.c_float), ("fcoeff", C.c_float), ("bcoeff", C.c_float),
                ("zcore", C.c_float), ("piewald", C.c_float), ("mode", C.c_int), ("zo_grad_bug", C.c_int),
                ("zmcore", C.c_float)] + [(k, C.c_float) for k in (
                    "zq2", "zq4", "zq22", "zq44", "zo2", "zo4", "zo6", "zo22", "zo44", "zo66",
                    "zh2", "zh4", "zh6", "zh8", "zh22", "zh44", "zh66", "zh88")]


# --electrostatic orders of ZD.mode (Config.cpp:114-128)
ZM_MODES = {"zero-dipole": 0, "zero-quadrupole": 1, "zero-octupole": 2, "zero-hexadecapole": 3}


class FF(C.Structure):
    """orc_ff_t: the HPS / Debye-Hueckel options of the pair form (the two
    tables are kept alive on the Python object)."""
    _fields_ = [("hps", C.c_int), ("dh", C.c_int), ("dh_grad_fix", C.c_int), ("hps_eps", C.c_float),
                ("dielectric_inv", C.c_float), ("debye_inv", C.c_float), ("hps_cutoff", C.POINTER(C.c_float)),
                ("hps_lambda", C.POINTER(C.c_float))]


def ff_setup(hps_cutoff, hps_lambda, hps_eps, dielectric, ionic_strength, temperature, hps=True, dh=True,
             dh_grad_fix=False) -> FF:
    """orc_dh_setup + the tables: HPS LJ switch and Debye-Hueckel electrostatics."""
    ff = FF()
    ff._hc = np.ascontiguousarray(hps_cutoff, np.float32).reshape(-1)
    ff._hl = np.ascontiguousarray(hps_lambda, np.float32).reshape(-1)
    ff.hps, ff.dh, ff.dh_grad_fix = int(hps), int(dh), int(dh_grad_fix)
    ff.hps_eps = float(np.float32(hps_eps))
    ff.hps_cutoff = ff._hc.ctypes.data_as(C.POINTER(C.c_float))
    ff.hps_lambda = ff._hl.ctypes.data_as(C.POINTER(C.c_float))
    if dh:
        lib().orc_dh_setup(C.byref(ff), C.c_double(dielectric), C.c_double(ionic_strength), C.c_double(temperature))
    return ff


def pair_ff_fp32(zd, ff, d12, t, a6, a12, q1, q2):
    d = np.ascontiguousarray(d12, np.float32)
    ev, ee = C.c_float(), C.c_float()
    w = np.zeros(3, np.float32)
    n = lib().orc_pair_ff_fp32(C.byref(zd), C.byref(ff) if ff is not None else None, _p(d, C.c_float), C.c_int(t),
                               C.c_float(a6), C.c_float(a12), C.c_float(q1), C.c_float(q2), C.byref(ev),
                               C.byref(ee), _p(w, C.c_float))
    return n, ev.value, ee.value, w


def allpairs_ff(zd, ff, L, crd, atom_type, lj6, lj12, charge, nb15off_table):
    """orc_pairs_allpairs_ff: the all-pairs loop with the pair form of ff."""
    n = len(crd)
    work = np.zeros((n, 3))
    en = np.zeros(2)
    f = lib().orc_pairs_allpairs_ff
    f.restype = C.c_int64
    npair = f(C.byref(zd), C.byref(ff), _p(np.ascontiguousarray(L, np.float32), C.c_float), C.c_int(n),
              _p(np.ascontiguousarray(crd, np.float64), C.c_double),
              _p(np.ascontiguousarray(atom_type, np.int32), C.c_int), C.c_int(lj6.shape[0]),
              _p(np.ascontiguousarray(lj6, np.float32), C.c_float),
              _p(np.ascontiguousarray(lj12, np.float32), C.c_float),
              _p(np.ascontiguousarray(charge, np.float32), C.c_float),
              _p(np.ascontiguousarray(nb15off_table, np.int32), C.c_int), _p(work, C.c_double), _p(en, C.c_double))
    return work, en, npair


def build():
    """Compile liboracle.so (always) and, where the reference tree is mounted,
    oracle/_ref/ from the reference sources."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/src/celeste/core"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
        _LIB = C.CDLL(path)
    return _LIB


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def grid_setup(n_atoms, box, nscut) -> Grid:
    g = Grid()
    rc = lib().orc_grid_setup(C.byref(g), C.c_int(n_atoms), _p(np.ascontiguousarray(box, np.float64), C.c_double),
                              C.c_double(nscut))
    if rc != 0:
        raise ValueError("orc_grid_setup failed (%d): box too small for the list radius" % rc)
    return g


class CellsResult:
    """numpy copies of an orc_cells_t after orc_cells_build."""

    def __init__(self, g: Grid, c: Cells, n_atoms: int):
        self.n_cells = c.n_cells
        self.n_atom_array = c.n_atom_array
        na = c.n_atom_array
        self.atomids = np.ctypeslib.as_array(c.atomids, (na,)).copy()
        self.crd = np.ctypeslib.as_array(c.crd, (na * 3,)).reshape(na, 3).copy()
        self.cell_crd = np.ctypeslib.as_array(c.cell_crd, (max(c.n_cells, 1) * 3,)).reshape(-1, 3)[:c.n_cells].copy()
        self.idx_xy_head_cell = np.ctypeslib.as_array(c.idx_xy_head_cell, (g.n_columns + 1,)).copy()
        self.idx_xy_head_atom = np.ctypeslib.as_array(c.idx_xy_head_atom, (g.n_columns + 1,)).copy()
        self.n_atoms_col = np.ctypeslib.as_array(c.n_atoms_col, (g.n_columns,)).copy()
        self.n_cells_z = np.ctypeslib.as_array(c.n_cells_z, (g.n_columns,)).copy()
        self.column_of_atom = np.ctypeslib.as_array(c.column_of_atom, (n_atoms,)).copy()


class Oracle:
    """One system's grid state in the oracle (owns the C buffers)."""

    def __init__(self, n_atoms, box, nscut):
        self.n_atoms = n_atoms
        self.g = grid_setup(n_atoms, box, nscut)
        self.c = Cells()
        lib().orc_cells_alloc(C.byref(self.c), C.byref(self.g), C.c_int(n_atoms))

    def __del__(self):
        try:
            lib().orc_cells_free(C.byref(self.c))
        except Exception:
            pass

    def build(self, crd):
        crd = np.ascontiguousarray(crd, np.float64)
        lib().orc_cells_build(C.byref(self.c), C.byref(self.g), C.c_int(self.n_atoms), _p(crd, C.c_double))
        return CellsResult(self.g, self.c, self.n_atoms)

    def enumerate(self, nb15off_table, gap_bug=True, cap=None):
        if cap is None:
            cap = max(self.g.max_n_cell_pairs, 1024)
        out = np.zeros(cap, CELLPAIR_DTYPE)
        head = np.zeros(self.c.n_cells + 1, np.int32)
        tab = np.ascontiguousarray(nb15off_table, np.int32)
        n = lib().orc_enumerate_cell_pairs(C.byref(self.g), C.byref(self.c), _p(tab, C.c_int), C.c_int(int(gap_bug)),
                                           out.ctypes.data_as(C.c_void_p), C.c_int(cap), _p(head, C.c_int))
        if n < 0:
            raise RuntimeError("cell-pair capacity exceeded")
        return out[:n].copy(), head

    def pairs_fp32(self, zd, cps, atom_type, lj6, lj12, charge):
        n_types = lj6.shape[0]
        work = np.zeros((self.c.n_atom_array, 3))
        en = np.zeros(2)
        cps = np.ascontiguousarray(cps)
        f = lib().orc_pairs_celllist
        f.restype = C.c_int64
        npair = f(C.byref(zd), C.byref(self.g), C.byref(self.c), cps.ctypes.data_as(C.c_void_p), C.c_int(len(cps)),
                  _p(np.ascontiguousarray(atom_type, np.int32), C.c_int), C.c_int(n_types),
                  _p(np.ascontiguousarray(lj6, np.float32), C.c_float),
                  _p(np.ascontiguousarray(lj12, np.float32), C.c_float),
                  _p(np.ascontiguousarray(charge, np.float32), C.c_float), _p(work, C.c_double), _p(en, C.c_double))
        return work, en, npair

    def pairs_ff(self, zd, ff, cps, atom_type, lj6, lj12, charge):
        """orc_pairs_celllist_ff: the cell-list loop with the pair form of ff (grid-order work)."""
        work = np.zeros((self.c.n_atom_array, 3))
        en = np.zeros(2)
        cps = np.ascontiguousarray(cps)
        f = lib().orc_pairs_celllist_ff
        f.restype = C.c_int64
        npair = f(C.byref(zd), C.byref(ff), C.byref(self.g), C.byref(self.c), cps.ctypes.data_as(C.c_void_p),
                  C.c_int(len(cps)), _p(np.ascontiguousarray(atom_type, np.int32), C.c_int), C.c_int(lj6.shape[0]),
                  _p(np.ascontiguousarray(lj6, np.float32), C.c_float),
                  _p(np.ascontiguousarray(lj12, np.float32), C.c_float),
                  _p(np.ascontiguousarray(charge, np.float32), C.c_float), _p(work, C.c_double), _p(en, C.c_double))
        return work, en, npair

    def pairs_fp64(self, zd, cps, crd, atom_type, lj6, lj12, charge):
        n_types = lj6.shape[0]
        work = np.zeros((self.n_atoms, 3))
        en = np.zeros(3)  # vdw, ele, virial
        cps = np.ascontiguousarray(cps)
        f = lib().orc_pairs_celllist_fp64
        f.restype = C.c_int64
        npair = f(C.byref(zd), C.byref(self.g), C.byref(self.c), cps.ctypes.data_as(C.c_void_p), C.c_int(len(cps)),
                  _p(np.ascontiguousarray(crd, np.float64), C.c_double),
                  _p(np.ascontiguousarray(atom_type, np.int32), C.c_int), C.c_int(n_types),
                  _p(np.ascontiguousarray(lj6, np.float32), C.c_float),
                  _p(np.ascontiguousarray(lj12, np.float32), C.c_float),
                  _p(np.ascontiguousarray(charge, np.float32), C.c_float), _p(work, C.c_double), _p(en, C.c_double))
        return work, en, npair


def zd_setup(cutoff, alpha, mode=0, zo_grad_bug=True) -> ZD:
    """orc_zm_setup: the constants of zero-multipole order `mode` (ZM_MODES)."""
    z = ZD()
    lib().orc_zm_setup(C.byref(z), C.c_double(cutoff), C.c_double(alpha), C.c_int(int(mode)),
                       C.c_int(int(zo_grad_bug)))
    return z


def self_energy(zd, charge, excess_pairs):
    q = np.ascontiguousarray(charge, np.float32)
    ex = np.ascontiguousarray(excess_pairs, np.int32)
    e = np.zeros(len(q))
    f = lib().orc_self_energy
    f.restype = C.c_double
    tot = f(C.byref(zd), C.c_int(len(q)), _p(q, C.c_float), C.c_int(len(ex)), _p(ex, C.c_int), _p(e, C.c_double))
    return tot, e


def pair_fp32(zd, d12, a6, a12, q1, q2):
    d = np.ascontiguousarray(d12, np.float32)
    ev, ee = C.c_float(), C.c_float()
    w = np.zeros(3, np.float32)
    n = lib().orc_pair_fp32(C.byref(zd), _p(d, C.c_float), C.c_float(a6), C.c_float(a12), C.c_float(q1),
                            C.c_float(q2), C.byref(ev), C.byref(ee), _p(w, C.c_float))
    return n, ev.value, ee.value, w


def allpairs_fp32(zd, L, crd, atom_type, lj6, lj12, charge, nb15off_table):
    n = len(crd)
    work = np.zeros((n, 3))
    en = np.zeros(2)
    Lf = np.ascontiguousarray(L, np.float32)
    f = lib().orc_pairs_allpairs_fp32
    f.restype = C.c_int64
    npair = f(C.byref(zd), _p(Lf, C.c_float), C.c_int(n), _p(np.ascontiguousarray(crd, np.float64), C.c_double),
              _p(np.ascontiguousarray(atom_type, np.int32), C.c_int), C.c_int(lj6.shape[0]),
              _p(np.ascontiguousarray(lj6, np.float32), C.c_float),
              _p(np.ascontiguousarray(lj12, np.float32), C.c_float),
              _p(np.ascontiguousarray(charge, np.float32), C.c_float),
              _p(np.ascontiguousarray(nb15off_table, np.int32), C.c_int), _p(work, C.c_double), _p(en, C.c_double))
    return work, en, npair


def excess(zd, L, pairs, crd, charge, work, bug_compat=True):
    f = lib().orc_excess
    f.restype = C.c_double
    return f(C.byref(zd), _p(np.ascontiguousarray(L, np.float32), C.c_float), C.c_int(len(pairs)),
             _p(np.ascontiguousarray(pairs, np.int32), C.c_int), _p(np.ascontiguousarray(crd, np.float64), C.c_double),
             _p(np.ascontiguousarray(charge, np.float32), C.c_float), C.c_int(int(bug_compat)), _p(work, C.c_double))


def bonds(L, ids, k, r0, crd, work):
    f = lib().orc_bonds
    f.restype = C.c_double
    return f(_p(np.ascontiguousarray(L, np.float32), C.c_float), C.c_int(len(ids)),
             _p(np.ascontiguousarray(ids, np.int32), C.c_int), _p(np.ascontiguousarray(k, np.float64), C.c_double),
             _p(np.ascontiguousarray(r0, np.float64), C.c_double), _p(np.ascontiguousarray(crd, np.float64), C.c_double),
             _p(work, C.c_double))


def pbc_from_box(box):
    L = (C.c_float * 3)()
    lb = (C.c_float * 3)()
    ub = (C.c_float * 3)()
    lib().orc_pbc_from_box(_p(np.ascontiguousarray(box, np.float64), C.c_double), L, lb, ub)
    return np.array(L[:], np.float32), np.array(lb[:], np.float32), np.array(ub[:], np.float32)


# ---- integration wrappers ----

def update_velocities(dt, vel, work, mass_inv):
    out = np.zeros_like(vel)
    lib().orc_update_velocities(C.c_int(len(vel)), C.c_double(dt), _p(np.ascontiguousarray(vel), C.c_double),
                                _p(np.ascontiguousarray(work), C.c_double),
                                _p(np.ascontiguousarray(mass_inv, np.float32), C.c_float), _p(out, C.c_double))
    return out


def update_coordinates(dt, crd_prev, vel):
    out = np.zeros_like(crd_prev)
    lib().orc_update_coordinates(C.c_int(len(vel)), C.c_double(dt), _p(np.ascontiguousarray(crd_prev), C.c_double),
                                 _p(np.ascontiguousarray(vel), C.c_double), _p(out, C.c_double))
    return out


def shake_trio(L, ids, dist2, crd, crd_prev, mass_inv, tol, max_loops):
    crd = np.ascontiguousarray(crd).copy()
    nf = lib().orc_shake_trio(_p(np.ascontiguousarray(L, np.float32), C.c_float), C.c_int(len(ids)),
                              _p(np.ascontiguousarray(ids, np.int32), C.c_int),
                              _p(np.ascontiguousarray(dist2, np.float64), C.c_double), _p(crd, C.c_double),
                              _p(np.ascontiguousarray(crd_prev), C.c_double),
                              _p(np.ascontiguousarray(mass_inv, np.float32), C.c_float), C.c_double(tol),
                              C.c_int(max_loops))
    return crd, nf


def shake_pair(L, ids, dist2, crd, crd_prev, mass_inv):
    crd = np.ascontiguousarray(crd).copy()
    lib().orc_shake_pair(_p(np.ascontiguousarray(L, np.float32), C.c_float), C.c_int(len(ids)),
                         _p(np.ascontiguousarray(ids, np.int32), C.c_int),
                         _p(np.ascontiguousarray(dist2, np.float64), C.c_double), _p(crd, C.c_double),
                         _p(np.ascontiguousarray(crd_prev), C.c_double),
                         _p(np.ascontiguousarray(mass_inv, np.float32), C.c_float))
    return crd


def settle(L, ids, dOH, dHH, crd, crd_prev, mass):
    crd = np.ascontiguousarray(crd).copy()
    lib().orc_settle(_p(np.ascontiguousarray(L, np.float32), C.c_float), C.c_int(len(ids)),
                     _p(np.ascontiguousarray(ids, np.int32), C.c_int), C.c_double(dOH), C.c_double(dHH),
                     _p(crd, C.c_double), _p(np.ascontiguousarray(crd_prev), C.c_double),
                     _p(np.ascontiguousarray(mass, np.float32), C.c_float))
    return crd


def velocity_from_crd(dt, crd, crd_prev):
    out = np.zeros_like(crd)
    lib().orc_velocity_from_crd(C.c_int(len(crd)), C.c_double(1.0 / dt), _p(np.ascontiguousarray(crd), C.c_double),
                                _p(np.ascontiguousarray(crd_prev), C.c_double), _p(out, C.c_double))
    return out


def revise_pbc(L, lb, crd):
    crd = np.ascontiguousarray(crd).copy()
    lib().orc_revise_pbc(C.c_int(len(crd)), _p(np.ascontiguousarray(L, np.float32), C.c_float),
                         _p(np.ascontiguousarray(lb, np.float32), C.c_float), _p(crd, C.c_double))
    return crd


def kinetic(mass, vel, vel_next):
    f = lib().orc_kinetic
    f.restype = C.c_double
    return f(C.c_int(len(vel)), _p(np.ascontiguousarray(mass, np.float32), C.c_float),
             _p(np.ascontiguousarray(vel), C.c_double), _p(np.ascontiguousarray(vel_next), C.c_double))


def cancel_com(groups, mass, vel):
    """groups: list of atom-index arrays."""
    vel = np.ascontiguousarray(vel).copy()
    off = np.zeros(len(groups) + 1, np.int32)
    for i, g in enumerate(groups):
        off[i + 1] = off[i] + len(g)
    atoms = np.concatenate([np.asarray(g, np.int32) for g in groups]) if groups else np.zeros(0, np.int32)
    m32 = np.asarray(mass, np.float32)
    minv = np.zeros(len(groups), np.float32)
    for i, g in enumerate(groups):
        s = np.float32(0.0)
        for a in g:   # float accumulation like MmSystem.cpp:572-582
            s = np.float32(s + m32[a])
        minv[i] = np.float32(1.0 / np.float64(s))
    lib().orc_cancel_com(C.c_int(len(groups)), _p(off, C.c_int), _p(atoms, C.c_int), _p(minv, C.c_float),
                         _p(m32, C.c_float), _p(vel, C.c_double))
    return vel


def langevin_positions(dt_half, gamma, T, L, crd_prev, crd_prev2, work, mass, mass_inv, zeta):
    out = np.zeros_like(crd_prev)
    lib().orc_langevin_positions(C.c_int(len(out)), C.c_double(dt_half), C.c_double(gamma), C.c_double(T),
                                 _p(np.ascontiguousarray(L, np.float32), C.c_float),
                                 _p(np.ascontiguousarray(crd_prev), C.c_double),
                                 _p(np.ascontiguousarray(crd_prev2), C.c_double),
                                 _p(np.ascontiguousarray(work), C.c_double),
                                 _p(np.ascontiguousarray(mass, np.float32), C.c_float),
                                 _p(np.ascontiguousarray(mass_inv, np.float32), C.c_float),
                                 _p(np.ascontiguousarray(zeta), C.c_double), _p(out, C.c_double))
    return out


def langevin_velocity(dt, L, crd, crd_prev2):
    out = np.zeros_like(crd)
    lib().orc_langevin_velocity(C.c_int(len(crd)), C.c_double(dt), _p(np.ascontiguousarray(L, np.float32), C.c_float),
                                _p(np.ascontiguousarray(crd), C.c_double), _p(np.ascontiguousarray(crd_prev2), C.c_double),
                                _p(out, C.c_double))
    return out


def run_ref_minicell(n_atoms, box, nscut, crd, nb15off_table, workdir):
    """Run oracle/_ref/minicell_drv (the reference MiniCell compiled from its
    sources) and parse its dump.  Only where _ref was built."""
    exe = os.path.join(_HERE, "_ref", "minicell_drv")
    inp = os.path.join(workdir, "mc_in.bin")
    out = os.path.join(workdir, "mc_out.bin")
    with open(inp, "wb") as f:
        f.write(np.int32(n_atoms).tobytes())
        f.write(np.ascontiguousarray(box, np.float64).tobytes())
        f.write(np.float64(nscut).tobytes())
        f.write(np.int32(0).tobytes())
        f.write(np.ascontiguousarray(crd, np.float64).tobytes())
        f.write(np.ascontiguousarray(nb15off_table, np.int32).tobytes())
    subprocess.run([exe, inp, out], check=True, stdout=subprocess.DEVNULL)
    b = open(out, "rb").read()
    n_cells, n_cols, n_cp = np.frombuffer(b, np.int32, 3, 0)
    o = 12
    atomids = np.frombuffer(b, np.int32, n_cells * 8, o).copy()
    o += 4 * n_cells * 8
    gcrd = np.frombuffer(b, np.float64, n_cells * 24, o).reshape(-1, 3).copy()
    o += 8 * n_cells * 24
    head_xy = np.frombuffer(b, np.int32, n_cols + 1, o).copy()
    o += 4 * (n_cols + 1)
    head_cp = np.frombuffer(b, np.int32, n_cells + 1, o).copy()
    o += 4 * (n_cells + 1)
    cps = np.frombuffer(b, CELLPAIR_DTYPE, n_cp, o).copy()
    return dict(n_cells=int(n_cells), atomids=atomids, crd=gcrd, idx_xy_head_cell=head_xy,
                idx_head_cell_pairs=head_cp, cell_pairs=cps)


def vcmd_bias(L, atoms_a, atoms_b, crdg, crd, mass, n_states_lo, n_states_hi, vs, sigma, coef, T, work):
    """orc_vcmd_bias: VcMD 1-D bias into work (in place) and crdg (in place).
    Returns (energy, lambda, in_range, (cand0, cand1))."""
    f = lib().orc_vcmd_bias
    f.restype = C.c_double
    m32 = np.ascontiguousarray(mass, np.float32)
    a = np.ascontiguousarray(atoms_a, np.int32)
    b = np.ascontiguousarray(atoms_b, np.int32)
    ma = np.float32(m32[a].sum(dtype=np.float32))
    mb = np.float32(m32[b].sum(dtype=np.float32))
    minv_a = np.float32(1.0 / np.float64(ma))
    minv_b = np.float32(1.0 / np.float64(mb))
    lo = np.ascontiguousarray(n_states_lo, np.float64)
    hi = np.ascontiguousarray(n_states_hi, np.float64)
    lam = C.c_double()
    inr = C.c_int()
    cand = (C.c_int * 2)()
    assert crdg.dtype == np.float64 and crdg.flags.c_contiguous and work.dtype == np.float64
    e = f(_p(np.ascontiguousarray(L, np.float32), C.c_float), C.c_int(len(a)), _p(a, C.c_int), C.c_int(len(b)),
          _p(b, C.c_int), _p(crdg, C.c_double), _p(np.ascontiguousarray(crd, np.float64), C.c_double),
          _p(m32, C.c_float), C.c_float(minv_a), C.c_float(minv_b), C.c_int(len(lo)), _p(lo, C.c_double),
          _p(hi, C.c_double), C.c_int(vs), C.c_double(sigma), C.c_double(coef), C.c_double(T), _p(work, C.c_double),
          C.byref(lam), C.byref(inr), cand)
    return e, lam.value, bool(inr.value), (cand[0], cand[1])


def ref_normals(seed, skip, n):
    """orc_ref_normals: the reference's N(0,1) draws (std::mt19937 seeded with
    seed, a fresh std::normal_distribution per draw, Random.cpp:10-43)."""
    out = np.zeros(int(n))
    lib().orc_ref_normals(C.c_int(int(seed)), C.c_longlong(int(skip)), C.c_longlong(int(n)), _p(out, C.c_double))
    return out


JOULE_CAL = 4.184        # CelesteObject.cpp
GAS_CONST = 8.31451      # CelesteObject.cpp:38
KINETIC_COEFF = (1e7 / (4.184 * 1e3)) * 0.5


def temperature_coeff(d_free):
    """ThermostatObject::set_temperature_coeff (Thermostat.cpp:13-17)."""
    return (2.0 * JOULE_CAL * 1e+3) / (GAS_CONST * float(d_free)) * KINETIC_COEFF


def thermostat_tau_inv(dt, tau):
    """SubBox::init_thermostat (SubBox.cpp:1674-1676): tau = --berendsen-tau (fs),
    replaced by the time step when the step is the smaller."""
    if dt < tau:
        tau = dt
    return 1.0 / tau


def thermostat_scaling(dt, tau_inv, temperature, temp_coeff, work, vel, mass, mass_inv):
    """orc_thermostat_scaling (ThermostatScaling::apply_thermostat): returns (vel_next, scale)."""
    out = np.zeros_like(np.asarray(vel, np.float64))
    f = lib().orc_thermostat_scaling
    f.restype = C.c_double
    s = f(C.c_int(len(out)), C.c_double(dt), C.c_double(tau_inv), C.c_double(temperature), C.c_double(temp_coeff),
          _p(np.ascontiguousarray(work, np.float64), C.c_double), _p(np.ascontiguousarray(vel, np.float64), C.c_double),
          _p(np.ascontiguousarray(mass, np.float32), C.c_float),
          _p(np.ascontiguousarray(mass_inv, np.float32), C.c_float), _p(out, C.c_double))
    return out, s


def bonded(ls, crd, work):
    """orc_angles + orc_torsions (torsions, impropers) + orc_nb14 on atom-order
    crd, adding into work; returns the five energies."""
    t = ls.bonded_tables()
    L = np.ascontiguousarray(pbc_from_box(ls.box)[0], np.float32)
    crd = np.ascontiguousarray(crd, np.float64)
    lf = lib()
    lf.orc_angles.restype = C.c_double
    lf.orc_torsions.restype = C.c_double
    ids, k, th = t["angle"]
    ea = lf.orc_angles(_p(L, C.c_float), C.c_int(len(ids)), _p(np.ascontiguousarray(ids), C.c_int),
                       _p(np.ascontiguousarray(k), C.c_double), _p(np.ascontiguousarray(th), C.c_double),
                       _p(crd, C.c_double), _p(work, C.c_double))
    et = []
    for key in ("torsion", "impro"):
        ids, e, ov, sy, ph = t[key]
        et.append(lf.orc_torsions(_p(L, C.c_float), C.c_int(len(ids)), _p(np.ascontiguousarray(ids), C.c_int),
                                  _p(np.ascontiguousarray(e), C.c_double), _p(np.ascontiguousarray(ov), C.c_int),
                                  _p(np.ascontiguousarray(sy), C.c_int), _p(np.ascontiguousarray(ph), C.c_double),
                                  _p(crd, C.c_double), _p(work, C.c_double)))
    ids, ty, cv, ce = t["nb14"]
    lj6, lj12 = ls.lj_tables_f32()
    ev, ee = C.c_double(), C.c_double()
    lf.orc_nb14(_p(L, C.c_float), C.c_int(len(ids)), _p(np.ascontiguousarray(ids), C.c_int),
                _p(np.ascontiguousarray(ty), C.c_int), _p(np.ascontiguousarray(cv), C.c_double),
                _p(np.ascontiguousarray(ce), C.c_double), _p(np.ascontiguousarray(lj6), C.c_float),
                _p(np.ascontiguousarray(lj12), C.c_float), C.c_int(lj6.shape[0]),
                _p(np.ascontiguousarray(ls.charge, np.float32), C.c_float), _p(crd, C.c_double),
                _p(work, C.c_double), C.byref(ev), C.byref(ee))
    return {"angle": ea, "torsion": et[0], "impro": et[1], "vdw14": ev.value, "ele14": ee.value}


def shake_quad(L, ids, dist2, crd, crd_prev, mass_inv, tol, max_loops):
    """orc_shake_quad on a copy of crd; returns (crd, n_not_converged)."""
    out = np.ascontiguousarray(crd, np.float64).copy()
    n = lib().orc_shake_quad(_p(np.ascontiguousarray(L, np.float32), C.c_float), C.c_int(len(ids)),
                             _p(np.ascontiguousarray(ids, np.int32), C.c_int),
                             _p(np.ascontiguousarray(dist2, np.float64), C.c_double), _p(out, C.c_double),
                             _p(np.ascontiguousarray(crd_prev, np.float64), C.c_double),
                             _p(np.ascontiguousarray(mass_inv, np.float32), C.c_float), C.c_double(tol),
                             C.c_int(max_loops))
    return out, n


def _group_minv(m32, ids):
    """Float group mass accumulated in order and its float inverse (MmSystem.cpp:572-582)."""
    s = np.float32(0.0)
    for i in ids:
        s = np.float32(s + m32[i])
    return np.float32(1.0 / np.float64(s))


def vs_transition(lo, hi, p_down, p_up, vs, lam, u_up, u_down):
    """orc_vs_transition: the V-McMD / V-AUS state after a transition trial."""
    a = [np.ascontiguousarray(x, np.float64) for x in (lo, hi, p_down, p_up)]
    return lib().orc_vs_transition(C.c_int(len(a[0])), *[_p(x, C.c_double) for x in a], C.c_int(vs),
                                   C.c_double(lam), C.c_double(u_up), C.c_double(u_down))


def vmcmd_factor(poly, lo, hi, T, lam):
    """orc_vmcmd_factor: the V-McMD gradient factor of one state."""
    f = lib().orc_vmcmd_factor
    f.restype = C.c_double
    p = np.ascontiguousarray(poly, np.float64)
    return f(_p(p, C.c_double), C.c_int(len(p) - 1), C.c_double(lo), C.c_double(hi), C.c_double(T), C.c_double(lam))


def vaus_bias(L, atoms_a, atoms_b, crdg, crd, mass, poly, lo, hi, sigma, coef, T, work):
    """orc_vaus_bias: V-AUS bias of one state into work (in place), crdg updated
    in place; returns lambda."""
    f = lib().orc_vaus_bias
    f.restype = C.c_double
    m32 = np.ascontiguousarray(mass, np.float32)
    a = np.ascontiguousarray(atoms_a, np.int32)
    b = np.ascontiguousarray(atoms_b, np.int32)
    p = np.ascontiguousarray(poly, np.float64)
    assert crdg.dtype == np.float64 and crdg.flags.c_contiguous and work.dtype == np.float64
    return f(_p(np.ascontiguousarray(L, np.float32), C.c_float), C.c_int(len(a)), _p(a, C.c_int), C.c_int(len(b)),
             _p(b, C.c_int), _p(crdg, C.c_double), _p(np.ascontiguousarray(crd, np.float64), C.c_double),
             _p(m32, C.c_float), C.c_float(_group_minv(m32, a)), C.c_float(_group_minv(m32, b)), _p(p, C.c_double),
             C.c_int(len(p) - 1), C.c_double(lo), C.c_double(hi), C.c_double(sigma), C.c_double(coef), C.c_double(T),
             _p(work, C.c_double))
