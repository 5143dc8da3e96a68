"""Shared helpers for the test-suite (fixtures, oracle MD loop)."""
import os

import numpy as np

import oracle
from omegagene_amd.cls import LaunchSet, read_cls

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_npz(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def ala3_launchset():
    d = load_npz("ala3_step0.npz")
    ls = LaunchSet(box=d["box"], crd=d["crd"], vel=d["vel"], charge=d["charge"], mass=d["mass"],
                   atom_type=d["atom_type"], n_lj_types=d["lj6"].shape[0], lj6=d["lj6"], lj12=d["lj12"],
                   nb15off=d["nb15off"])
    expected = dict(zip([str(k) for k in d["expected_keys"]], d["expected_vals"]))
    return ls, expected


def wrap_inbox(crd, L, lb, ub):
    """MmSystem::revise_crd_inbox (MmSystem.cpp:808-816)."""
    crd = crd.copy()
    for d in range(3):
        while (crd[:, d] < lb[d]).any():
            crd[crd[:, d] < lb[d], d] += L[d]
        while (crd[:, d] >= ub[d]).any():
            crd[crd[:, d] >= ub[d], d] -= L[d]
    return crd


def cls_from_fixture(npz):
    return read_cls(bytes(npz["cls"]))


def rms_rel(a, b):
    return float(np.sqrt(np.mean(np.sum((a - b) ** 2, axis=1))) / np.sqrt(np.mean(np.sum(b ** 2, axis=1))))


def oracle_forces(ls: LaunchSet, cutoff=12.0, alpha=0.0, nscut=13.0, excess_bug=True, crd=None):
    """All reference force terms with the oracle: cell-list pairs (FP32 rules),
    excess, bonds; returns (work_atom, energies dict, oracle objects)."""
    crd = ls.crd if crd is None else crd
    zd = oracle.zd_setup(cutoff, alpha)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    o = oracle.Oracle(ls.n_atoms, ls.box, nscut)
    c = o.build(crd)
    tab = ls.nb15off_table()
    cps, head = o.enumerate(tab, gap_bug=True)
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    wg, en, npair = o.pairs_fp32(zd, cps, ls.atom_type, lj6, lj12, q)
    work = np.zeros((ls.n_atoms, 3))
    m = c.atomids >= 0
    work[c.atomids[m]] = wg[m]
    ex = ls.excess_pairs()
    eex = oracle.excess(zd, L, ex, crd, q, work, bug_compat=excess_bug)
    eb = oracle.bonds(L, ls.bond_ids, ls.bond_k, ls.bond_r0, crd, work) if len(ls.bond_ids) else 0.0
    eself, _ = oracle.self_energy(zd, q, ex)
    return work, dict(vdw=en[0], ele=en[1], excess=eex, bond=eb, self=eself, n_pairs=npair), (o, c, cps, head)


def ele_scale(ls: LaunchSet, cutoff=12.0, alpha=0.0, nscut=13.0, crd=None):
    """Sum over in-cutoff pairs of |q_i q_j| * 332.06378 * k(r), with
    k(r) = erfc(alpha r)/r - zcore + bcoeff r^2 >= 0 on (0, rc]: the zero-dipole
    pair energy with every charge replaced by its magnitude.  The net ele
    energy is a near-cancelling sum of such terms, so its FP32 rounding error
    scales with this sum, not with the net value."""
    crd = ls.crd if crd is None else crd
    zd = oracle.zd_setup(cutoff, alpha)
    o = oracle.Oracle(ls.n_atoms, ls.box, nscut)
    o.build(crd)
    cps, _ = o.enumerate(ls.nb15off_table(), gap_bug=True)
    lj6, lj12 = ls.lj_tables_f32()
    _, en, _ = o.pairs_fp32(zd, cps, ls.atom_type, lj6, lj12, np.abs(ls.charge).astype(np.float32))
    return float(en[1])


def oracle_forces_allpairs(ls: LaunchSet, crd, zd, L, alpha_bug=True, ff=None):
    """The reference all-pairs build's force terms (SubBox.cpp:1127-1220 pairs,
    :1347-1367 excess, :1229-1249 bonds) with the oracle: work [n][3] and the
    potential pieces (vdw, ele incl. excess and self, bond).  ff: the HPS /
    Debye-Hueckel form (oracle.ff_setup); under Debye-Hueckel the excluded
    pairs add nothing (calc_null, ZeroMultipoleSum.cpp:456-466) and the self
    energy is zero (zmcore = 0 outside the zero-multipole modes, :75-82)."""
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    if ff is not None:
        work, en, _ = oracle.allpairs_ff(zd, ff, L, crd, ls.atom_type, lj6, lj12, q, ls.nb15off_table())
        if ff.dh:
            eb = oracle.bonds(L, ls.bond_ids, ls.bond_k, ls.bond_r0, crd, work) if len(ls.bond_ids) else 0.0
            return work, en[0], en[1], eb
    else:
        work, en, _ = oracle.allpairs_fp32(zd, L, crd, ls.atom_type, lj6, lj12, q, ls.nb15off_table())
    ex = ls.excess_pairs()
    eex = oracle.excess(zd, L, ex, crd, q, work, bug_compat=alpha_bug)
    eb = oracle.bonds(L, ls.bond_ids, ls.bond_k, ls.bond_r0, crd, work) if len(ls.bond_ids) else 0.0
    eself, _ = oracle.self_energy(zd, q, ex)
    return work, en[0], en[1] + eex + eself, eb


def cfg_value(npz, key, default=None):
    """A key of the reference .cfg text stored in a fixture."""
    for line in bytes(npz["cfg"]).decode().splitlines():
        parts = line.split()
        if parts and parts[0] == "--" + key:
            return parts[1]
    return default


def hps_dh_form(ls: LaunchSet, npz, dh_grad_fix=False):
    """oracle.ff_setup from a fixture's .cfg keys (Config.cpp:327-342) and the launch set's HPS tables."""
    hc, hl = ls.hps_tables_f32()
    return oracle.ff_setup(hc, hl, float(cfg_value(npz, "hydrophobicity-scale-epsiron")),
                           float(cfg_value(npz, "debye-huckel-dielectric")),
                           float(cfg_value(npz, "debye-huckel-ionic-strength")),
                           float(cfg_value(npz, "debye-huckel-temperature")), dh_grad_fix=dh_grad_fix)


def oracle_md(ls: LaunchSet, n_steps, dt=2.0, cutoff=12.0, integrator="leapfrog-presto", constraint="none",
              tol=1e-10, max_loops=1000, com_groups=None, thermostat=None, langevin=None, extra_last=False,
              ff=None, zd=None, ext=None, ext_out=None):
    """The reference step loop with the oracle's functions, in the reference
    order:
      leapfrog-presto (DynamicsModePresto::calc_in_each_step, DynamicsMode.cpp:468-560):
        forces; vel <- vel_next; vel_next += dt a (SubBox.cpp:1445-1454); COM
        cancel (COMMotion.cpp:30-55); scaling thermostat when unconstrained
        (SubBox.cpp:1745-1761, Thermostat.cpp:57-84); x += dt vel_next;
        SHAKE pairs then trios (ConstraintShake.cpp:10-16) and
        v = dx/dt (SubBox.cpp:1739-1743); PBC wrap; KE of (vel + vel_next)/2;
      langevin (DynamicsModeLangevin::calc_in_each_step, DynamicsMode.cpp:798-899):
        forces; x_prev2 <- x_prev <- x; step 0: x += dt v; later steps the
        position Langevin update (SubBox.cpp:2077-2114) with the reference
        noise stream, v by central difference (SubBox.cpp:2054-2066) and its
        KE; no constraints (DynamicsMode.cpp:877); PBC wrap.
    ff: the HPS / Debye-Hueckel pair form (oracle.ff_setup), None for zero-multipole.
    zd: the zero-multipole constants (oracle.zd_setup); None: zero-dipole at alpha 0.
    ext: V-McMD / V-AUS on the gradient before the velocity update
    (DynamicsMode.cpp:504-507, Extend.cpp:126-540): dict(mode "v-mcmd" or
    "v-aus", vmcmd VmcmdDef, and for V-AUS groups (a, b), crdg [unwrapped
    group coordinates, updated], sigma, recov); transitions every interval
    steps with both uniform draws 0.5 (the probabilities in {0, 1} of the
    fixtures make the walk independent of the draws); ext_out collects
    (lambda, state) per step.
    thermostat: dict(tau_fs, temperature, temperature_init, heating_steps, d_free).
    langevin: dict(gamma_per_fs, temperature, seed).
    extra_last: the reference runs one more step after the last output
    (DynamicsMode.cpp:201-208); its restart then holds x_prev and v of that
    step for Langevin.  Returns (crd, vel, rows[step, pot, kin, bond, vdw, ele]),
    crd/vel being what the reference's restart holds."""
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    zd = oracle.zd_setup(cutoff, 0.0) if zd is None else zd
    m32 = ls.mass.astype(np.float32)
    minv = (1.0 / m32.astype(np.float64)).astype(np.float32)
    vel_next = ls.vel.copy()
    vel = vel_next.copy()
    crd_prev = crd.copy()
    crd_prev2 = crd.copy()
    rows = []
    n_total = n_steps + (1 if extra_last else 0)
    n = ls.n_atoms
    for step in range(n_total):
        work, evdw, eele, ebond = oracle_forces_allpairs(ls, crd, zd, L, ff=ff)
        pot = evdw + eele + ebond
        if len(ls.angles) or len(ls.torsions) or len(ls.impros) or len(ls.nb14):
            # angles, torsions, impropers, 1-4 (SubBox.cpp:1250-1346)
            pot += sum(oracle.bonded(ls, crd, work).values())
        if ext is not None:
            mc = ext["vmcmd"]
            st = mc.states
            lo = [x[1] for x in st]
            hi = [x[2] for x in st]
            if "vs" not in ext:
                ext["vs"] = mc.init_vs
            if ext["mode"] == "v-mcmd":
                lam = pot
            else:
                m_ = ls.mass
                lam = oracle.vaus_bias(L, ext["groups"][0], ext["groups"][1], ext["crdg"], crd, m_, [0.0], 0.0, 0.0,
                                       0.0, 0.0, mc.temperature, np.zeros_like(work))
            if step > 0 and step % mc.interval == 0:
                ext["vs"] = oracle.vs_transition(lo, hi, [x[3] for x in st], [x[4] for x in st], ext["vs"], lam,
                                                 0.5, 0.5)
            vs = ext["vs"]
            if ext["mode"] == "v-mcmd":
                work = work * oracle.vmcmd_factor(st[vs][5], lo[vs], hi[vs], mc.temperature, lam)
            else:
                oracle.vaus_bias(L, ext["groups"][0], ext["groups"][1], ext["crdg"], crd, ls.mass, st[vs][5], lo[vs],
                                 hi[vs], ext["sigma"], ext["recov"], mc.temperature, work)
            if ext_out is not None:
                ext_out.append((lam, vs))
        if integrator == "leapfrog-presto":
            vel = vel_next
            vel_next = oracle.update_velocities(dt, vel, work, minv)
            if com_groups:
                vel_next = oracle.cancel_com(com_groups, m32, vel_next)
            if thermostat and constraint == "none":
                t = thermostat
                T = t["temperature"]
                if step < t["heating_steps"]:
                    T = t["temperature_init"] + ((t["temperature"] - t["temperature_init"]) /
                                                 float(t["heating_steps"])) * step
                vel_next, _ = oracle.thermostat_scaling(dt, oracle.thermostat_tau_inv(dt, t["tau_fs"]), T,
                                                        oracle.temperature_coeff(t["d_free"]), work, vel, m32, minv)
            crd_prev = crd
            crd = oracle.update_coordinates(dt, crd_prev, vel_next)
            if constraint != "none":
                if len(ls.shake_pair_ids):
                    crd = oracle.shake_pair(L, ls.shake_pair_ids, np.asarray(ls.shake_pair_d) ** 2, crd, crd_prev, minv)
                if len(ls.shake_trio_ids):
                    crd, nf = oracle.shake_trio(L, ls.shake_trio_ids, np.asarray(ls.shake_trio_d) ** 2, crd, crd_prev,
                                                minv, tol, max_loops)
                    assert nf == 0
                if len(ls.shake_quads):
                    qids, qd2 = ls.shake_quad_tables()
                    crd, nf = oracle.shake_quad(L, qids, qd2, crd, crd_prev, minv, tol, max_loops)
                    assert nf == 0
                vel_next = oracle.velocity_from_crd(dt, crd, crd_prev)
            crd = oracle.revise_pbc(L, lb, crd)
            ke = oracle.kinetic(m32, vel, vel_next)
            out_crd, out_vel = crd, vel_next
        else:
            g = langevin
            crd_prev2, crd_prev = crd_prev, crd
            if step == 0:
                crd = oracle.update_coordinates(dt, crd_prev, vel_next)
                ke = np.nan
            else:
                zeta = oracle.ref_normals(g["seed"], (step - 1) * 3 * n, 3 * n).reshape(n, 3)
                crd = oracle.langevin_positions(dt / 2, g["gamma_per_fs"], g["temperature"], L, crd_prev, crd_prev2,
                                                work, m32, minv, zeta)
                vel = oracle.langevin_velocity(dt, L, crd, crd_prev2)
                ke = oracle.kinetic(m32, vel, vel)
            crd = oracle.revise_pbc(L, lb, crd)
            out_crd, out_vel = crd_prev, vel
        rows.append([step, pot, ke, ebond, evdw, eele])
    if integrator == "leapfrog-presto":
        return out_crd, out_vel, np.array(rows)
    return out_crd, out_vel, np.array(rows)


def chain_shake_pairs(ls: LaunchSet, every=3):
    """SHAKE pairs on every `every`-th bond of the chain (disjoint pairs), at the current bond lengths."""
    b = ls.bond_ids[::every]
    L = float(ls.box[0])
    d = ls.crd[b[:, 0]] - ls.crd[b[:, 1]]
    d -= L * np.round(d / L)
    ls.shake_pair_ids = b.astype(np.int32)
    ls.shake_pair_d = np.linalg.norm(d, axis=1)
    return ls
