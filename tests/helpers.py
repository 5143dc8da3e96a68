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
