"""The reference's GPU call points (include/og_celeste.h) from Python.

celeste's host (SubBox / MiniCell) drives its CUDA backend through a fixed
sequence of extern "C" calls (SubBox::set_nsgrid SubBox.cpp:487-545,
gpu_device_setup :1683-1711, update_device_cell_info :1713-1723,
nsgrid_crd_to_gpu :552-558, calc_energy :906-955).  ``CelesteGpu`` binds the
same calls on libog_engine.so under the og_ prefix, with the same argument
meaning; ``SubBoxDriver`` plays that sequence for one system whose grid is
built on the host, as MiniCell does.  Every call raises ``EngineError`` on a
non-zero status (the reference exits instead).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import EngineError, _check, lib

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)
_D = C.POINTER(C.c_double)


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def pbc_from_box(box):
    """Box lengths and lower bounds as SubBox hands them to og_set_pbc: float
    L from the diagonal of the 3x3 box, lb = centre - L/2 (PBC::set_pbc, PBC.cpp:18-37)."""
    box = np.asarray(box, np.float64)
    L = np.array([box[0], box[4], box[8]], np.float32)
    lb = np.array([np.float64(box[9 + i]) - np.float64(np.float32(L[i] * np.float32(0.5))) for i in range(3)],
                  np.float32)
    return L, lb


class ZeroDipole:
    """ZD constants at alpha = 0 in the float arithmetic of
    ZeroMultipoleSum::set_zms_params (ZeroMultipoleSum.cpp:108-190): what
    SubBox passes to og_zerodipole_constant."""

    def __init__(self, cutoff):
        c = np.float32(cutoff)
        c2 = np.float32(c * c)
        c3 = np.float32(c2 * c)
        self.fcoeff = np.float32(1.0 / np.float64(c3))
        self.bcoeff = np.float32(0.5 * np.float64(self.fcoeff))
        self.zcore = np.float32(np.float32(np.float32(1.0) / c) + np.float32(self.bcoeff * c2))


class CelesteGpu:
    """Pinned host buffers and the call points, named as in cuda_setup.cu."""

    def __init__(self):
        self.L = lib()
        self.h_crd = _F()
        self.h_atomids = _I()
        self.h_work = _D()
        self.h_energy = _D()
        self.h_head = _I()
        self.h_type = _I()
        self.h_charge = _F()
        self.max_n_atom_array = 0
        self.n_columns = 0
        self.n_atoms_system = 0

    # ---- host allocation (MiniCell.cpp:82-87, SubBox.cpp:124) ----
    def hostalloc_atom_info(self, max_n_atom_array):
        _check(self.L.og_hostalloc_atom_info(C.byref(self.h_crd), C.byref(self.h_atomids), C.byref(self.h_work),
                                             C.byref(self.h_energy), C.c_int(max_n_atom_array)))
        self.max_n_atom_array = max_n_atom_array

    def hostalloc_cell_info(self, n_columns):
        _check(self.L.og_hostalloc_cell_info(C.byref(self.h_head), C.c_int(n_columns)))
        self.n_columns = n_columns

    def hostalloc_atom_type_charge(self, n_atoms_system):
        _check(self.L.og_hostalloc_atom_type_charge(C.byref(self.h_type), C.byref(self.h_charge),
                                                    C.c_int(n_atoms_system)))
        self.n_atoms_system = n_atoms_system

    # ---- device setup (SubBox::gpu_device_setup, SubBox.cpp:1683-1711) ----
    def alloc_atom_info(self, max_n_atoms_exbox, max_n_cells, max_n_cell_pairs, n_columns):
        _check(self.L.og_alloc_atom_info(C.c_int(max_n_atoms_exbox), C.c_int(max_n_cells),
                                         C.c_int(max_n_cell_pairs), C.c_int(n_columns)))

    def alloc_set_lj_params(self, lj6, lj12, nb15off, max_n_nb15off):
        lj6 = np.ascontiguousarray(lj6, np.float32)
        lj12 = np.ascontiguousarray(lj12, np.float32)
        tab = np.ascontiguousarray(nb15off, np.int32)
        # SubBox.cpp:1693-1694 passes the two sizes after max_n_nb15off as well
        _check(self.L.og_alloc_set_lj_params(_p(lj6, C.c_float), _p(lj12, C.c_float), C.c_int(lj6.shape[0]),
                                             _p(tab, C.c_int), C.c_int(max_n_nb15off), C.c_int(tab.shape[0]),
                                             C.c_int(0)))

    def set_pbc(self, L, lb):
        _check(self.L.og_set_pbc(_p(np.ascontiguousarray(L, np.float32), C.c_float),
                                 _p(np.ascontiguousarray(lb, np.float32), C.c_float)))

    def set_constant(self, cutoff, cutoff_pairlist, n_atomtypes):
        _check(self.L.og_set_constant(C.c_float(cutoff), C.c_float(cutoff_pairlist), C.c_int(n_atomtypes)))

    def zerodipole_constant(self, zcore, bcoeff, fcoeff):
        _check(self.L.og_zerodipole_constant(C.c_float(zcore), C.c_float(bcoeff), C.c_float(fcoeff)))

    def memcpy_htod_atom_info(self, charge, atom_type):
        n = self.n_atoms_system
        np.ctypeslib.as_array(self.h_charge, (n,))[:] = np.asarray(charge, np.float32)
        np.ctypeslib.as_array(self.h_type, (n,))[:] = np.asarray(atom_type, np.int32)
        _check(self.L.og_memcpy_htod_atom_info(C.byref(self.h_charge), C.byref(self.h_type)))

    # ---- per rebuild (SubBox::update_device_cell_info, SubBox.cpp:1713-1723) ----
    def set_cell_constant(self, n_cells, n_atoms_exbox, n_atom_array, n_cells_xyz, n_columns, l_cell_xyz,
                          n_neighbor_xyz):
        nc = np.ascontiguousarray(n_cells_xyz, np.int32)
        lc = np.ascontiguousarray(l_cell_xyz, np.float32)
        nn = np.ascontiguousarray(n_neighbor_xyz, np.int32)
        _check(self.L.og_set_cell_constant(C.c_int(n_cells), C.c_int(n_atoms_exbox), C.c_int(n_atom_array),
                                           _p(nc, C.c_int), C.c_int(n_columns), _p(lc, C.c_float), _p(nn, C.c_int)))

    def memcpy_htod_atomids(self, atomids, idx_xy_head_cell):
        np.ctypeslib.as_array(self.h_atomids, (len(atomids),))[:] = atomids
        np.ctypeslib.as_array(self.h_head, (len(idx_xy_head_cell),))[:] = idx_xy_head_cell
        _check(self.L.og_memcpy_htod_atomids(C.byref(self.h_atomids), C.byref(self.h_head)))

    def set_atominfo(self):
        _check(self.L.og_set_atominfo())

    # ---- per step (SubBox::nsgrid_crd_to_gpu, SubBox.cpp:552-558) ----
    def memcpy_htod_crd(self, grid_crd_f32):
        a = np.ascontiguousarray(grid_crd_f32, np.float32).reshape(-1)
        np.ctypeslib.as_array(self.h_crd, (len(a),))[:] = a
        _check(self.L.og_memcpy_htod_crd(C.byref(self.h_crd)))

    def set_crd(self):
        _check(self.L.og_set_crd())

    def enumerate_cell_pairs(self, n_cells, n_neighbor_col, idx_atom_cell_xy):
        col = np.ascontiguousarray(idx_atom_cell_xy, np.int32)
        _check(self.L.og_enumerate_cell_pairs(C.byref(self.h_atomids), C.c_int(n_cells), C.c_int(n_neighbor_col),
                                              _p(col, C.c_int)))

    def pairwise_ljzd(self, flg_mod_15mask=False):
        _check(self.L.og_pairwise_ljzd(C.c_bool(bool(flg_mod_15mask))))

    # ---- the HPS + Debye-Hueckel form (SubBox.cpp:1702-1710,1730-1733) ----
    def hps_constant(self, eps):
        _check(self.L.og_hps_constant(C.c_float(eps)))

    def alloc_set_hps_params(self, hps_cutoff, hps_lambda):
        hc = np.ascontiguousarray(hps_cutoff, np.float32)
        hl = np.ascontiguousarray(hps_lambda, np.float32)
        _check(self.L.og_alloc_set_hps_params(_p(hc, C.c_float), _p(hl, C.c_float), C.c_int(hc.shape[0])))

    def debye_huckel_constant(self, dielect, temperature, ionic_strength):
        _check(self.L.og_debye_huckel_constant(C.c_float(dielect), C.c_float(temperature), C.c_float(ionic_strength)))

    def pairwise_hps_dh(self, flg_mod_15mask=False):
        _check(self.L.og_pairwise_hps_dh(C.c_bool(bool(flg_mod_15mask))))

    def pair_sync(self):
        _check(self.L.og_pair_sync())

    def memcpy_dtoh_work(self, n_atoms, n_atom_array):
        _check(self.L.og_memcpy_dtoh_work(C.byref(self.h_work), C.byref(self.h_energy), C.c_int(n_atoms),
                                          C.c_int(n_atom_array)))
        w = np.ctypeslib.as_array(self.h_work, (n_atom_array * 3,)).reshape(n_atom_array, 3).copy()
        e = np.ctypeslib.as_array(self.h_energy, (2,)).copy()
        return w, e

    def cell_pairs(self, n_cells):
        """The device cell-pair list as CellPair records (MiniCell.h:17-26) and its per-cell heads."""
        n = C.c_int(0)
        _check(self.L.og_get_celeste_cell_pairs(None, None, C.byref(n)))
        head = np.zeros(n_cells + 1, np.int32)
        cps = np.zeros((max(n.value, 1), 5), np.int32)
        _check(self.L.og_get_celeste_cell_pairs(_p(cps, C.c_int), _p(head, C.c_int), None))
        return cps[: n.value], head

    # ---- teardown (SubBox.cpp:91-92, MiniCell.cpp:189-193) ----
    def free(self):
        self.L.og_free_lj_params()
        self.L.og_free_atom_info()
        if self.h_crd:
            # work and energy by reference (real_fc *&, MiniCell.cpp:19): freed and nulled
            _check(self.L.og_hostfree_atom_info(self.h_crd, self.h_atomids, C.byref(self.h_work),
                                                C.byref(self.h_energy)))
            self.h_crd = _F()
        if self.h_head:
            _check(self.L.og_hostfree_cell_info(self.h_head))
            self.h_head = _I()
        if self.h_type:
            _check(self.L.og_hostfree_atom_type_charge(self.h_type, self.h_charge))
            self.h_type = _I()


class SubBoxDriver:
    """SubBox's GPU call sequence for one box whose MiniCell grid is given.

    ``grid`` supplies what MiniCell owns: n_cells, atomids (grid order),
    crd (grid-order coordinates, float-exact), idx_xy_head_cell,
    column_of_slot; ``geom`` the MiniCell geometry (n_cells_xyz, L_cell_xyz,
    n_neighbors_xyz, n_columns, max_n_cells, max_n_cell_pairs)."""

    def __init__(self, n_atoms, L, lb, cutoff, nscut, zd, lj6, lj12, nb15off, charge, atom_type, geom,
                 max_n_atom_array):
        self.gpu = CelesteGpu()
        self.n_atoms = n_atoms
        self.geom = geom
        g = self.gpu
        g.hostalloc_atom_type_charge(n_atoms)                                  # SubBox.cpp:124
        g.hostalloc_atom_info(max_n_atom_array)                                # MiniCell.cpp:82
        g.hostalloc_cell_info(geom["n_columns"])                               # MiniCell.cpp:84-87
        g.alloc_atom_info(n_atoms, geom["max_n_cells"], geom["max_n_cell_pairs"], geom["n_columns"] + 1)
        g.alloc_set_lj_params(lj6, lj12, nb15off, nb15off.shape[1])
        g.set_pbc(L, lb)
        g.set_constant(cutoff, nscut, lj6.shape[0])
        g.zerodipole_constant(zd.zcore, zd.bcoeff, zd.fcoeff)
        g.memcpy_htod_atom_info(charge, atom_type)                              # SubBox.cpp:516

    def update_grid(self, grid):
        """SubBox::update_device_cell_info + nsgrid_crd_to_gpu + cuda_enumerate_cell_pairs."""
        g, geo = self.gpu, self.geom
        n_atom_array = grid["n_cells"] * 8
        g.set_cell_constant(grid["n_cells"], self.n_atoms, n_atom_array, geo["n_cells_xyz"], geo["n_columns"],
                            geo["L_cell_xyz"], geo["n_neighbors_xyz"])
        g.memcpy_htod_atomids(grid["atomids"][:n_atom_array], grid["idx_xy_head_cell"])
        g.set_atominfo()
        self.set_coordinates(grid["crd"][:n_atom_array])
        nn = geo["n_neighbors_xyz"]
        g.enumerate_cell_pairs(grid["n_cells"], (2 * nn[0] + 1) * (2 * nn[1] + 1), grid["column_of_slot"])
        self.n_atom_array = n_atom_array

    def set_coordinates(self, grid_crd):
        self.gpu.memcpy_htod_crd(np.asarray(grid_crd, np.float32))
        self.gpu.set_crd()

    def set_hps_dh(self, hps_cutoff, hps_lambda, hps_eps, dielect, temperature, ionic_strength):
        """SubBox::gpu_device_setup of the HPS + Debye-Hueckel build (SubBox.cpp:1702-1710): later
        pair_forces calls evaluate that form (calc_energy_pairwise_cuda, SubBox.cpp:1724-1736)."""
        g = self.gpu
        g.alloc_set_hps_params(hps_cutoff, hps_lambda)
        g.hps_constant(hps_eps)
        g.debye_huckel_constant(dielect, temperature, ionic_strength)
        self.hps_dh = True

    def pair_forces(self):
        """calc_energy_pairwise_cuda + cuda_pair_sync + cuda_memcpy_dtoh_work."""
        if getattr(self, "hps_dh", False):
            self.gpu.pairwise_hps_dh(False)
            self.gpu.pair_sync()
            return self.gpu.memcpy_dtoh_work(self.n_atoms, self.n_atom_array)
        self.gpu.pairwise_ljzd(False)
        self.gpu.pair_sync()
        return self.gpu.memcpy_dtoh_work(self.n_atoms, self.n_atom_array)

    def close(self):
        self.gpu.L.og_free_hps_params()
        self.gpu.free()


OG_NOALLOC_MAX_ATOM_ARRAY = 262144  # include/og_celeste.h


def pairwise_ljzd_noalloc(xyzq, atom_type, lj_pair, n_lj_types, idx_head_cell_pairs, cell_pair_c2, cell_pair_mask,
                          n_cells, L, cutoff, zd, work, energy, flags=0, stream=0):
    """og_pairwise_ljzd_noalloc: the pair evaluation of pairwise_ljzd +
    memcpy_dtoh_work on caller-owned device buffers, given as device
    addresses (ints), asynchronous on ``stream`` (a hipStream_t as int).
    No counterpart in the reference; it allocates nothing on the device."""
    lbox = np.ascontiguousarray(L, np.float32)
    vp = C.c_void_p
    _check(lib().og_pairwise_ljzd_noalloc(
        vp(xyzq), vp(atom_type), vp(lj_pair), C.c_int(n_lj_types), vp(idx_head_cell_pairs), vp(cell_pair_c2),
        vp(cell_pair_mask), C.c_int(n_cells), _p(lbox, C.c_float), C.c_float(cutoff), C.c_float(zd.zcore),
        C.c_float(zd.bcoeff), C.c_float(zd.fcoeff), vp(work), vp(energy), vp(flags or None), vp(stream or None)))
