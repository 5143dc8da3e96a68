"""Python handle on libog_engine.so (the C ABI of include/og_engine.h).

This is host plumbing for tests, the bench and the serving loop: every
computation runs in the HIP library.  Loading fails loudly when the library
is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from .cls import LaunchSet

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libog_engine.so")
_LIB = None

OG_OK, OG_E_ARG, OG_E_CAPACITY, OG_E_NAN, OG_E_NOCONV, OG_E_DEVICE, OG_E_STATE = range(7)
TERMS = ["vdw", "ele", "excess", "self", "bond", "kinetic", "virial", "extend", "npairs", "lambda", "vstate",
         "angle", "torsion", "impro", "vdw14", "ele14"]
NTERMS = len(TERMS)
# the potential-energy terms (MmSystem's pote_* sum the reference prints as "Potential", DynamicsMode.cpp:272-300)
POTENTIAL_TERMS = ["vdw", "ele", "excess", "self", "bond", "angle", "torsion", "impro", "vdw14", "ele14"]
_POT_COLS = [TERMS.index(t) for t in POTENTIAL_TERMS]


def potential(rows):
    """Potential energy of energy-log rows (or one row / an energies dict)."""
    if isinstance(rows, dict):
        return sum(rows[t] for t in POTENTIAL_TERMS)
    r = np.asarray(rows)
    return r[..., _POT_COLS].sum(axis=-1)
INTEGRATORS = {"leapfrog-presto": 0, "langevin": 1}
CONSTRAINTS = {"none": 0, "shake": 1, "shake-settle": 2}
THERMOSTATS = {"none": 0, "scaling": 1}
# pair-kernel block orders (og_set_pair_shape_order / og_batch_set_pair_shape)
ORDERS = ("system", "round-robin", "xcd", "xcd-parts")
NONBONDS = {"lennard-jones": 0, "hydrophobicity-scale-lj": 1}     # Config.cpp:333-339
ELECTROSTATICS = {"zero-dipole": 0, "debye-huckel": 1, "zero-quadrupole": 2, "zero-octupole": 3,
                  "zero-hexadecapole": 4}                           # Config.cpp:114-128


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("og_engine status %d: %s" % (code, msg))
        self.code = code


class Desc(C.Structure):
    _fields_ = [("n_atoms", C.c_int), ("box", C.POINTER(C.c_double)), ("crd", C.POINTER(C.c_double)),
                ("vel", C.POINTER(C.c_double)), ("charge", C.POINTER(C.c_double)), ("mass", C.POINTER(C.c_double)),
                ("atom_type", C.POINTER(C.c_int)), ("n_lj_types", C.c_int), ("lj6", C.POINTER(C.c_double)),
                ("lj12", C.POINTER(C.c_double)), ("nb15off", C.POINTER(C.c_int)), ("n_excess", C.c_int),
                ("excess", C.POINTER(C.c_int)), ("n_bonds", C.c_int), ("bond_ids", C.POINTER(C.c_int)),
                ("bond_k", C.POINTER(C.c_double)), ("bond_r0", C.POINTER(C.c_double)),
                ("n_shake_pairs", C.c_int), ("shake_pair_ids", C.POINTER(C.c_int)),
                ("shake_pair_d2", C.POINTER(C.c_double)), ("n_shake_trios", C.c_int),
                ("shake_trio_ids", C.POINTER(C.c_int)), ("shake_trio_d2", C.POINTER(C.c_double)),
                ("n_settle", C.c_int), ("settle_ids", C.POINTER(C.c_int)), ("settle_dOH", C.c_double),
                ("settle_dHH", C.c_double), ("n_com_groups", C.c_int), ("com_group_off", C.POINTER(C.c_int)),
                ("com_group_atoms", C.POINTER(C.c_int)), ("d_free", C.c_int),
                ("vc_n_states", C.c_int), ("vc_min", C.POINTER(C.c_double)), ("vc_max", C.POINTER(C.c_double)),
                ("vc_q_cano", C.POINTER(C.c_double)), ("vc_init_state", C.c_int), ("vc_interval", C.c_int),
                ("vc_n_a", C.c_int), ("vc_n_b", C.c_int), ("vc_atoms_a", C.POINTER(C.c_int)),
                ("vc_atoms_b", C.POINTER(C.c_int)), ("hps_cutoff", C.POINTER(C.c_double)),
                ("hps_lambda", C.POINTER(C.c_double)),
                ("n_angles", C.c_int), ("angle_ids", C.POINTER(C.c_int)), ("angle_k", C.POINTER(C.c_double)),
                ("angle_theta0", C.POINTER(C.c_double)),
                ("n_torsions", C.c_int), ("torsion_ids", C.POINTER(C.c_int)), ("torsion_ene", C.POINTER(C.c_double)),
                ("torsion_overlaps", C.POINTER(C.c_int)), ("torsion_symmetry", C.POINTER(C.c_int)),
                ("torsion_phase", C.POINTER(C.c_double)),
                ("n_impros", C.c_int), ("impro_ids", C.POINTER(C.c_int)), ("impro_ene", C.POINTER(C.c_double)),
                ("impro_overlaps", C.POINTER(C.c_int)), ("impro_symmetry", C.POINTER(C.c_int)),
                ("impro_phase", C.POINTER(C.c_double)),
                ("n_nb14", C.c_int), ("nb14_ids", C.POINTER(C.c_int)), ("nb14_types", C.POINTER(C.c_int)),
                ("nb14_coeff_vdw", C.POINTER(C.c_double)), ("nb14_coeff_ele", C.POINTER(C.c_double)),
                ("n_shake_quads", C.c_int), ("shake_quad_ids", C.POINTER(C.c_int)),
                ("shake_quad_d2", C.POINTER(C.c_double)),
                ("mc_mode", C.c_int), ("mc_n_states", C.c_int), ("mc_interval", C.c_int),
                ("mc_temperature", C.c_double), ("mc_order", C.POINTER(C.c_int)), ("mc_poly", C.POINTER(C.c_double)),
                ("mc_lo", C.POINTER(C.c_double)), ("mc_hi", C.POINTER(C.c_double)),
                ("mc_prob_down", C.POINTER(C.c_double)), ("mc_prob_up", C.POINTER(C.c_double)),
                ("mc_init_state", C.c_int), ("vc_crd_init", C.POINTER(C.c_double))]


OG_EXT_VMCMD, OG_EXT_VAUS = 2, 3
OG_MC_MAX_ORDER = 8


def ext_mode(cfg: dict, ls: LaunchSet) -> str:
    """The --extended-ensemble of cfg (Config.cpp:145-156) when the launch set
    carries its block: "vcmd", "v-mcmd", "v-aus" or "none"."""
    m = cfg.get("extended-ensemble", "none")
    if m == "vcmd" and ls.vcmd is not None:
        return m
    if m in ("v-mcmd", "v-aus") and ls.vmcmd is not None and (m == "v-mcmd" or ls.group_coord is not None):
        return m
    return "none"


class Params(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("nsgrid_cutoff", C.c_double), ("nsgrid_update_intvl", C.c_int),
                ("ele_alpha", C.c_double), ("time_step", C.c_double), ("integrator", C.c_int),
                ("constraint", C.c_int), ("const_tolerance", C.c_double), ("const_max_loops", C.c_int),
                ("temperature", C.c_double), ("langevin_gamma", C.c_double), ("random_seed", C.c_uint64),
                ("com_cancel", C.c_int), ("gap_bug", C.c_int), ("excess_bug", C.c_int), ("log_stride", C.c_int),
                ("log_cap", C.c_int), ("capacity_factor", C.c_double), ("enhance_sigma", C.c_double),
                ("enhance_recov_coef", C.c_double), ("split_step", C.c_int), ("thermostat", C.c_int),
                ("berendsen_tau", C.c_double), ("temperature_init", C.c_double), ("heating_steps", C.c_int),
                ("nonbond", C.c_int), ("hps_epsilon", C.c_double), ("electrostatic", C.c_int),
                ("dh_dielectric", C.c_double), ("dh_ionic_strength", C.c_double), ("dh_temperature", C.c_double),
                ("dh_grad_fix", C.c_int), ("cell_edge_scale", C.c_double), ("zo_grad_bug", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libog_engine.so is not built (run python -m omegagene_amd.build); "
                               "the engine has no CPU fallback")
        _LIB = C.CDLL(LIB_PATH)
        _LIB.og_status_string.restype = C.c_char_p
    return _LIB


def _check(code):
    if code != OG_OK:
        buf = C.create_string_buffer(512)
        lib().og_last_error(buf, 512)
        raise EngineError(code, buf.value.decode() or lib().og_status_string(code).decode())


def params_from_cfg(cfg: dict) -> Params:
    """Map the reference .cfg keys (Config.cpp:37-367) onto og_params."""
    p = Params()
    lib().og_params_default(C.byref(p))
    g = lambda k, d=None: cfg.get(k, d)  # noqa: E731
    p.cutoff = float(g("cutoff", 12.0))
    p.nsgrid_cutoff = float(g("nsgrid-cutoff", p.cutoff + 1.0))
    p.nsgrid_update_intvl = int(g("nsgrid-update-intvl", 1))
    p.ele_alpha = float(g("ele-alpha", 0.0))
    p.time_step = float(g("time-step", 0.0005))
    p.integrator = INTEGRATORS[g("integrator", "leapfrog-presto")]
    p.constraint = CONSTRAINTS[g("constraint", "none")]
    p.const_tolerance = float(g("const-tolerance", 1e-6))
    p.const_max_loops = int(g("const-max-loops", 1000))
    p.temperature = float(g("temperature", 300.0))
    p.langevin_gamma = float(g("langevin-gamma", 0.0)) / 1000.0  # ps^-1 -> fs^-1 (Config.cpp:345-348)
    p.random_seed = int(g("random-seed", 0)) & 0xFFFFFFFFFFFFFFFF
    p.com_cancel = 1 if g("com-motion", "none") == "cancel" else 0
    for k, attr in (("gap-bug", "gap_bug"), ("excess-bug", "excess_bug"), ("log-stride", "log_stride"),
                    ("log-cap", "log_cap"), ("split-step", "split_step")):
        if k in cfg:
            setattr(p, attr, int(cfg[k]))
    if "capacity-factor" in cfg:
        p.capacity_factor = float(cfg["capacity-factor"])
    p.enhance_sigma = float(g("enhance-sigma", 0.0))                # Config.cpp --enhance-sigma
    p.enhance_recov_coef = float(g("enhance-recovery-coef", 0.0))   # Config.cpp --enhance-recovery-coef
    p.thermostat = THERMOSTATS[g("thermostat", "none")]             # Config.cpp:133-143
    p.berendsen_tau = float(g("berendsen-tau", 0.0)) * 1000.0       # ps -> fs (Config.cpp:162-164)
    p.temperature_init = float(g("temperature-init", -1.0))         # Config.cpp:166-167
    p.heating_steps = int(g("heating-steps", 0))                    # Config.cpp:169-170
    p.nonbond = NONBONDS[g("nonbond", "lennard-jones")]              # Config.cpp:333-339
    p.hps_epsilon = float(g("hydrophobicity-scale-epsiron", 0.0))   # Config.cpp:340-342
    p.electrostatic = ELECTROSTATICS[g("electrostatic", "zero-dipole")]
    p.dh_dielectric = float(g("debye-huckel-dielectric", 0.0))     # Config.cpp:327-332
    p.dh_ionic_strength = float(g("debye-huckel-ionic-strength", 0.0))
    p.dh_temperature = float(g("debye-huckel-temperature", 0.0))
    if "dh-grad-fix" in cfg:
        p.dh_grad_fix = int(cfg["dh-grad-fix"])
    if "cell-edge-scale" in cfg:
        p.cell_edge_scale = float(cfg["cell-edge-scale"])
    if "zo-grad-bug" in cfg:
        p.zo_grad_bug = int(cfg["zo-grad-bug"])
    return p


class _Keep:
    """Holds the numpy arrays a Desc points into."""


def desc_from_launchset(ls: LaunchSet, constraint: int = 1, com_groups=None, vcmd: bool = False, ext: str = "none"):
    """og_system_desc of a launch set; ext (or vcmd=True for "vcmd") selects
    the extended ensemble whose block the launch set carries."""
    if vcmd:
        ext = "vcmd"
    k = _Keep()

    def d(a, dt=np.float64):
        a = np.ascontiguousarray(a, dt)
        k.__dict__["a%d" % len(k.__dict__)] = a
        return a.ctypes.data_as(C.POINTER(C.c_double if dt == np.float64 else C.c_int))

    n = ls.n_atoms
    x = Desc()
    x.n_atoms = n
    x.box = d(ls.box)
    x.crd = d(ls.crd)
    x.vel = d(ls.vel)
    x.charge = d(ls.charge)
    x.mass = d(ls.mass)
    x.atom_type = d(ls.atom_type, np.int32)
    x.n_lj_types = ls.n_lj_types
    x.lj6 = d(ls.lj6)
    x.lj12 = d(ls.lj12)
    x.nb15off = d(ls.nb15off_table(), np.int32)
    ex = ls.excess_pairs()
    x.n_excess = len(ex)
    x.excess = d(ex, np.int32)
    x.n_bonds = len(ls.bond_ids)
    x.bond_ids = d(ls.bond_ids, np.int32)
    x.bond_k = d(ls.bond_k)
    x.bond_r0 = d(ls.bond_r0)
    x.n_shake_pairs = len(ls.shake_pair_ids)
    x.shake_pair_ids = d(ls.shake_pair_ids, np.int32)
    x.shake_pair_d2 = d(np.asarray(ls.shake_pair_d) ** 2)
    x.n_shake_trios = len(ls.shake_trio_ids)
    x.shake_trio_ids = d(ls.shake_trio_ids, np.int32)
    x.shake_trio_d2 = d(np.asarray(ls.shake_trio_d) ** 2)
    x.n_settle = len(ls.settle_ids)
    x.settle_ids = d(ls.settle_ids, np.int32)
    if len(ls.settle_ids):
        x.settle_dOH = float(ls.settle_d[0][0])
        x.settle_dHH = float(ls.settle_d[0][1])
    groups = com_groups if com_groups is not None else [g for _, g in ls.groups]
    off = np.zeros(len(groups) + 1, np.int32)
    for i, g_ in enumerate(groups):
        off[i + 1] = off[i] + len(g_)
    x.n_com_groups = len(groups)
    x.com_group_off = d(off, np.int32)
    x.com_group_atoms = d(np.concatenate([np.asarray(g_, np.int32) for g_ in groups]) if groups else
                          np.zeros(1, np.int32), np.int32)
    x.d_free = ls.d_free(constraint != 0)
    if ls.hps_cutoff is not None:
        x.hps_cutoff = d(ls.hps_cutoff)
        x.hps_lambda = d(ls.hps_lambda)
    # angles, torsions, impropers, 1-4 pairs and four-atom SHAKE groups of the launch set
    bt = ls.bonded_tables()
    ids, kk, th = bt["angle"]
    if len(ids):
        x.n_angles, x.angle_ids, x.angle_k, x.angle_theta0 = len(ids), d(ids, np.int32), d(kk), d(th)
    for key in ("torsion", "impro"):
        ids, e, ov, sy, ph = bt[key]
        if len(ids):
            setattr(x, "n_%ss" % key, len(ids))
            for name, arr, dt in (("ids", ids, np.int32), ("ene", e, np.float64), ("overlaps", ov, np.int32),
                                  ("symmetry", sy, np.int32), ("phase", ph, np.float64)):
                setattr(x, "%s_%s" % (key, name), d(arr, dt))
    ids, ty, cv, ce = bt["nb14"]
    if len(ids):
        x.n_nb14, x.nb14_ids, x.nb14_types = len(ids), d(ids, np.int32), d(ty, np.int32)
        x.nb14_coeff_vdw, x.nb14_coeff_ele = d(cv), d(ce)
    qids, qd2 = ls.shake_quad_tables()
    if len(qids):
        x.n_shake_quads, x.shake_quad_ids, x.shake_quad_d2 = len(qids), d(qids, np.int32), d(qd2)
    vc = ls.vcmd if ext == "vcmd" else None
    if vc is not None:
        # 1-D VcMD on the distance between the mass centres of the dimension's two groups (Read.cpp:705-779)
        gids, _names, ranges = vc.dims[0]
        ga = np.asarray(ls.groups[gids[0] - 1][1], np.int32)
        gb = np.asarray(ls.groups[gids[1] - 1][1], np.int32)
        q = np.ones(len(ranges))
        for st, w in vc.q_cano:
            q[st[0]] = w
        x.vc_n_states = len(ranges)
        x.vc_min = d(np.array([r[0] for r in ranges]))
        x.vc_max = d(np.array([r[1] for r in ranges]))
        x.vc_q_cano = d(q)
        x.vc_init_state = int(vc.init_vs[0])
        x.vc_interval = int(vc.interval)
        x.vc_n_a, x.vc_n_b = len(ga), len(gb)
        x.vc_atoms_a = d(ga, np.int32)
        x.vc_atoms_b = d(gb, np.int32)
        gc = ls.group_coord
        if gc is not None and list(gc.group_ids[:2]) == list(gids[:2]):
            x.vc_crd_init = d(np.concatenate(gc.crd[:2]))   # the launch set's unwrapped group coordinates
    if ext in ("v-mcmd", "v-aus"):
        # V-McMD / V-AUS block (Read.cpp:521-561); V-AUS groups from the group-coordinate block (:646-703)
        mc = ls.vmcmd
        ns = len(mc.states)
        poly = np.zeros((ns, OG_MC_MAX_ORDER + 1))
        for i, st in enumerate(mc.states):
            if st[0] > OG_MC_MAX_ORDER:
                raise ValueError("virtual state %d: polynomial order %d above %d" % (i, st[0], OG_MC_MAX_ORDER))
            poly[i, :st[0] + 1] = st[5][:st[0] + 1]
        x.mc_mode = OG_EXT_VMCMD if ext == "v-mcmd" else OG_EXT_VAUS
        x.mc_n_states, x.mc_interval, x.mc_temperature = ns, int(mc.interval), float(mc.temperature)
        x.mc_order = d(np.array([st[0] for st in mc.states]), np.int32)
        x.mc_poly = d(poly)
        x.mc_lo, x.mc_hi = d(np.array([st[1] for st in mc.states])), d(np.array([st[2] for st in mc.states]))
        x.mc_prob_down = d(np.array([st[3] for st in mc.states]))
        x.mc_prob_up = d(np.array([st[4] for st in mc.states]))
        x.mc_init_state = int(mc.init_vs)
        if ext == "v-aus":
            gc = ls.group_coord
            ga = np.asarray(ls.groups[gc.group_ids[0] - 1][1], np.int32)
            gb = np.asarray(ls.groups[gc.group_ids[1] - 1][1], np.int32)
            x.vc_n_a, x.vc_n_b = len(ga), len(gb)
            x.vc_atoms_a, x.vc_atoms_b = d(ga, np.int32), d(gb, np.int32)
            x.vc_crd_init = d(np.concatenate(gc.crd[:2]))
    return x, k


def autotune_cell_edge(ls: LaunchSet, cfg: dict, trials: int = 5, com_groups=None):
    """og_autotune_cell_edge: the fastest MiniCell column-width scale for this
    system and configuration (1.0 under the reference list rule, gap-bug 1)
    and the pair-kernel time it gave (ms)."""
    _check(lib().og_set_device(C.c_int(0)))
    p = params_from_cfg(cfg)
    x, keep = desc_from_launchset(ls, p.constraint, com_groups, vcmd=False)
    sc, ms = C.c_double(), C.c_double()
    _check(lib().og_autotune_cell_edge(C.byref(x), C.byref(p), C.c_int(trials), C.byref(sc), C.byref(ms)))
    del keep
    return sc.value, ms.value


class System:
    """One simulated system resident on the GPU (og_system)."""

    def __init__(self, ls: LaunchSet, cfg: dict, device: int = 0, com_groups=None):
        L = lib()
        _check(L.og_set_device(C.c_int(device)))
        self.params = params_from_cfg(cfg)
        self.desc, self._keep = desc_from_launchset(ls, self.params.constraint, com_groups, ext=ext_mode(cfg, ls))
        self.n_atoms = ls.n_atoms
        self.d_free = self.desc.d_free
        self.h = C.c_void_p()
        _check(L.og_system_create(C.byref(self.desc), C.byref(self.params), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().og_system_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nsgrid_update(self):
        _check(lib().og_nsgrid_update(self.h))

    def energy_force(self, rebuild=True):
        w = np.zeros((self.n_atoms, 3))
        e = np.zeros(NTERMS)
        _check(lib().og_calc_energy_force(self.h, C.c_int(int(rebuild)), w.ctypes.data_as(C.c_void_p),
                                          e.ctypes.data_as(C.c_void_p)))
        return w, dict(zip(TERMS, e))

    def run(self, n_steps: int):
        _check(lib().og_md_run(self.h, C.c_int(n_steps)))

    def state(self):
        crd = np.zeros((self.n_atoms, 3))
        vn = np.zeros((self.n_atoms, 3))
        vj = np.zeros((self.n_atoms, 3))
        _check(lib().og_get_state(self.h, crd.ctypes.data_as(C.c_void_p), vn.ctypes.data_as(C.c_void_p),
                                  vj.ctypes.data_as(C.c_void_p)))
        return crd, vn, vj

    def set_state(self, crd=None, vel_next=None):
        c = np.ascontiguousarray(crd, np.float64) if crd is not None else None
        v = np.ascontiguousarray(vel_next, np.float64) if vel_next is not None else None
        _check(lib().og_set_state(self.h, c.ctypes.data_as(C.c_void_p) if c is not None else None,
                                  v.ctypes.data_as(C.c_void_p) if v is not None else None))

    def energies(self):
        e = np.zeros(NTERMS)
        _check(lib().og_get_energies(self.h, e.ctypes.data_as(C.c_void_p)))
        return dict(zip(TERMS, e))

    def energy_log(self, max_rows=100000):
        out = np.zeros((max_rows, NTERMS))
        n = C.c_int()
        _check(lib().og_get_energy_log(self.h, out.ctypes.data_as(C.c_void_p), C.c_int(max_rows), C.byref(n)))
        return out[:n.value]

    def step(self):
        s = C.c_longlong()
        _check(lib().og_get_step(self.h, C.byref(s)))
        return s.value

    def vcmd(self, n_states: int):
        """Current virtual state and visit counts per state."""
        vs = C.c_int()
        cnt = np.zeros(n_states)
        _check(lib().og_get_vcmd(self.h, C.byref(vs), cnt.ctypes.data_as(C.c_void_p), C.c_int(n_states)))
        return vs.value, cnt

    def grid_info(self):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().og_get_grid_info(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return dict(n_cells=a.value, n_columns=b.value, n_cell_pairs=c.value, n_atom_array=d.value)

    def grid(self):
        gi = self.grid_info()
        ids = np.zeros(gi["n_atom_array"], np.int32)
        gcrd = np.zeros((gi["n_atom_array"], 3))
        head = np.zeros(gi["n_columns"] + 1, np.int32)
        _check(lib().og_get_grid(self.h, ids.ctypes.data_as(C.c_void_p), gcrd.ctypes.data_as(C.c_void_p),
                                 head.ctypes.data_as(C.c_void_p)))
        return ids, gcrd, head

    def grid_geometry(self):
        """MiniCell geometry in the arguments of the call points (og_set_cell_constant, og_alloc_atom_info)."""
        nc = np.zeros(3, np.int32)
        lc = np.zeros(3, np.float32)
        nn = np.zeros(3, np.int32)
        mc, mp = C.c_int(), C.c_int()
        _check(lib().og_get_grid_geometry(self.h, nc.ctypes.data_as(C.c_void_p), lc.ctypes.data_as(C.c_void_p),
                                          nn.ctypes.data_as(C.c_void_p), C.byref(mc), C.byref(mp)))
        return {"n_cells_xyz": nc, "L_cell_xyz": lc, "n_neighbors_xyz": nn, "n_columns": int(nc[0] * nc[1]),
                "max_n_cells": mc.value, "max_n_cell_pairs": mp.value}

    def cell_pairs(self):
        gi = self.grid_info()
        cp = np.zeros((gi["n_cell_pairs"], 5), np.int32)
        head = np.zeros(gi["n_cells"] + 1, np.int32)
        _check(lib().og_get_cell_pairs(self.h, cp.ctypes.data_as(C.c_void_p), head.ctypes.data_as(C.c_void_p)))
        return cp, head

    def work_grid(self):
        gi = self.grid_info()
        w = np.zeros((gi["n_atom_array"], 3))
        _check(lib().og_get_work_grid(self.h, w.ctypes.data_as(C.c_void_p)))
        return w

    def set_pair_timing(self, on=True):
        _check(lib().og_set_pair_timing(self.h, C.c_int(int(on))))

    def status(self):
        """og_status of the system's last run or evaluation (0 = ok)."""
        st = C.c_int()
        _check(lib().og_system_status(self.h, C.byref(st)))
        return st.value

    def set_langevin_noise(self, zeta):
        """N(0,1) draws [n][3] for the next step's Langevin update (None cancels)."""
        if zeta is None:
            _check(lib().og_set_langevin_noise(self.h, None))
            return
        z = np.ascontiguousarray(zeta, np.float64).reshape(-1)
        assert z.size == 3 * self.n_atoms
        _check(lib().og_set_langevin_noise(self.h, z.ctypes.data_as(C.POINTER(C.c_double))))

    def work(self):
        """Gradient dE/dx of the last evaluation or step, [n][3] atom order."""
        w = np.zeros((self.n_atoms, 3))
        _check(lib().og_get_work(self.h, w.ctypes.data_as(C.POINTER(C.c_double))))
        return w

    def autotune(self, trials=5):
        """Time the pair kernel over its launch shapes and keep the fastest
        (results do not depend on it); returns the chosen shape and time."""
        w, ms, order = C.c_int(), C.c_double(), C.c_int()
        _check(lib().og_autotune(self.h, C.c_int(trials), C.byref(w), C.byref(ms)))
        _check(lib().og_get_pair_shape(self.h, C.byref(w), C.byref(order)))
        return {"waves_per_block": w.value, "order": ORDERS[order.value], "pair_ms": ms.value}

    def set_pair_shape(self, waves, order="system"):
        _check(lib().og_set_pair_shape_order(self.h, C.c_int(waves), C.c_int(ORDERS.index(order))))

    def device_bytes(self):
        b = C.c_uint64()
        _check(lib().og_system_device_bytes(self.h, C.byref(b)))
        return b.value

    def step_fused(self):
        """True when an MD step runs as the fused one-kernel Langevin step."""
        f = C.c_int()
        _check(lib().og_step_fused(self.h, C.byref(f)))
        return bool(f.value)

    def pair_time(self):
        ms, n = C.c_double(), C.c_int()
        _check(lib().og_pair_kernel_time(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Batch:
    """Many systems advanced together by the same launches (og_batch)."""

    def __init__(self):
        self.h = C.c_void_p()
        _check(lib().og_batch_create(C.byref(self.h)))
        self.members = []

    def add(self, s: System):
        _check(lib().og_batch_add(self.h, s.h))
        self.members.append(s)

    def remove(self, s: System):
        _check(lib().og_batch_remove(self.h, s.h))
        self.members.remove(s)

    def autotune(self, trials=5):
        """Time the pair kernel over its launch shapes (waves per block, block
        order across systems and XCDs) and keep the fastest; results do not
        depend on it."""
        w, order, ms = C.c_int(), C.c_int(), C.c_double()
        _check(lib().og_batch_autotune(self.h, C.c_int(trials), C.byref(w), C.byref(order), C.byref(ms)))
        return {"waves_per_block": w.value, "order": ORDERS[order.value], "pair_ms": ms.value}

    def set_pair_shape(self, waves, order=0):
        """order: 0 / "system" system order, 1 / "round-robin" across systems, 2 / "xcd" each system on one
        XCD, 3 / "xcd-parts" the list in eight contiguous parts, one per XCD."""
        o = ORDERS.index(order) if isinstance(order, str) else int(order)
        _check(lib().og_batch_set_pair_shape(self.h, C.c_int(waves), C.c_int(o)))

    def set_graph_steps(self, steps: int):
        """Steps per captured hipGraph (0: automatic)."""
        _check(lib().og_batch_set_graph_steps(self.h, C.c_int(steps)))

    def run(self, n_steps: int):
        _check(lib().og_batch_run(self.h, C.c_int(n_steps)))

    def close(self):
        if self.h:
            lib().og_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def footprint(ls: LaunchSet, cfg: dict) -> int:
    p = params_from_cfg(cfg)
    d, keep = desc_from_launchset(ls, p.constraint, ext=ext_mode(cfg, ls))
    b = C.c_uint64()
    _check(lib().og_system_footprint(C.byref(d), C.byref(p), C.byref(b)))
    return int(b.value)


def exports(header: str = None):
    """Names of the C-ABI symbols declared in include/*.h (or one header)."""
    import glob
    import re
    inc = os.path.join(os.path.dirname(_HERE), "include")
    paths = [os.path.join(inc, header)] if header else sorted(glob.glob(os.path.join(inc, "*.h")))
    names = set()
    for p in paths:
        hdr = open(p).read()
        names |= set(re.findall(r"^\s*(?:og_status|int|void|const char \*)\s*(og_\w+)\s*\(", hdr, re.M))
    return sorted(names)


class Server:
    """Continuous-batching segment service (og_server_*): walkers stay on the
    device; requests are admitted against an HBM budget and advanced
    together, chunk_steps per pump."""

    def __init__(self, hbm_budget_bytes: int, chunk_steps: int = 10, device: int = 0):
        _check(lib().og_set_device(C.c_int(device)))
        self.h = C.c_void_p()
        _check(lib().og_server_create(C.c_uint64(int(hbm_budget_bytes)), C.c_int(chunk_steps), C.byref(self.h)))
        self._n_atoms = {}   # walker -> atoms
        self._req_walker = {}
        self._req_steps = {}

    def submit_new(self, ls: LaunchSet, cfg: dict, n_steps: int):
        p = params_from_cfg(cfg)
        d, keep = desc_from_launchset(ls, p.constraint, ext=ext_mode(cfg, ls))
        rid, wid = C.c_longlong(), C.c_longlong()
        _check(lib().og_server_submit_new(self.h, C.byref(d), C.byref(p), C.c_int(n_steps), C.byref(rid),
                                          C.byref(wid)))
        self._n_atoms[wid.value] = ls.n_atoms
        self._req_walker[rid.value] = wid.value
        self._req_steps[rid.value] = n_steps
        return rid.value, wid.value

    def submit_next(self, walker: int, n_steps: int):
        rid = C.c_longlong()
        _check(lib().og_server_submit_next(self.h, C.c_longlong(walker), C.c_int(n_steps), C.byref(rid)))
        self._req_walker[rid.value] = walker
        self._req_steps[rid.value] = n_steps
        return rid.value

    def pump(self):
        a, q, d = C.c_int(), C.c_int(), C.c_int()
        _check(lib().og_server_pump(self.h, C.byref(a), C.byref(q), C.byref(d)))
        return a.value, q.value, d.value

    def result(self, rid: int):
        n = self._n_atoms[self._req_walker[rid]]
        steps = self._req_steps[rid]
        crd = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        rows = np.zeros((steps, NTERMS))
        nr = C.c_int()
        t = np.zeros(3)
        _check(lib().og_server_result(self.h, C.c_longlong(rid), crd.ctypes.data_as(C.c_void_p),
                                      vel.ctypes.data_as(C.c_void_p), rows.ctypes.data_as(C.c_void_p),
                                      C.c_int(steps), C.byref(nr), t.ctypes.data_as(C.c_void_p)))
        return crd, vel, rows[:nr.value], t

    def release_walker(self, walker: int):
        _check(lib().og_server_release_walker(self.h, C.c_longlong(walker)))

    def stats(self):
        t, s, b, w = C.c_double(), C.c_longlong(), C.c_uint64(), C.c_int()
        _check(lib().og_server_stats(self.h, C.byref(t), C.byref(s), C.byref(b), C.byref(w)))
        return dict(t=t.value, steps_done=s.value, bytes_used=b.value, n_walkers=w.value)

    def close(self):
        if self.h:
            lib().og_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
