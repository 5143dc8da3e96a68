"""The hydrophobicity-scale LJ + Debye-Hueckel pair form on the device
(kernel_pairwise_hps_dh / cal_pair_hps_dh, cuda_setup.cu:1273-1477; CPU
ForceField.cpp:333-418, ZeroMultipoleSum.cpp:443-455) against the oracle and
against the reference all-pairs build's own runs of coarse-grained chains
(tests/golden/wons_hpsdh_*.npz).

Bars: gradient RMS relative error <= 2e-6 against the FP32-faithful oracle on
the same cell-pair list and against the reference's all-pairs gradient; vdw
within 2e-6 relative; ele within 2e-6 of the sum of its term magnitudes (the
device takes exp(-kappa r) from the hardware exp, the reference from expf)."""
import sys

import numpy as np
import pytest

import oracle
from helpers import cfg_value, cls_from_fixture, hps_dh_form, load_npz, rms_rel, wrap_inbox
from omegagene_amd.celeste import SubBoxDriver
from omegagene_amd.engine import System

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _cfg(d, **kw):
    keys = ("cutoff", "nsgrid-cutoff", "electrostatic", "debye-huckel-dielectric", "debye-huckel-temperature",
            "debye-huckel-ionic-strength", "nonbond", "hydrophobicity-scale-epsiron", "time-step")
    c = {k: cfg_value(d, k) for k in keys}
    for k in ("cutoff", "nsgrid-cutoff", "debye-huckel-dielectric", "debye-huckel-temperature",
              "debye-huckel-ionic-strength", "hydrophobicity-scale-epsiron", "time-step"):
        c[k] = float(c[k])
    c.update({"nsgrid-update-intvl": 1, "integrator": "leapfrog-presto", "constraint": "none", "log-cap": 64})
    c.update(kw)
    return c


def _oracle_celllist(ls, crd, cfg, ff, gap_bug):
    """Cell-list pairs (FP32-faithful) + bonds in atom order, the energies and
    the sum of the ele term magnitudes."""
    zd = oracle.zd_setup(cfg["cutoff"], 0.0)
    o = oracle.Oracle(ls.n_atoms, ls.box, cfg["nsgrid-cutoff"])
    c = o.build(crd)
    cps, _ = o.enumerate(ls.nb15off_table(), gap_bug=gap_bug)
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    wg, en, npair = o.pairs_ff(zd, ff, cps, ls.atom_type, lj6, lj12, q)
    _, en_abs, _ = o.pairs_ff(zd, ff, cps, ls.atom_type, lj6, lj12, np.abs(q))
    w = np.zeros((ls.n_atoms, 3))
    m = c.atomids >= 0
    w[c.atomids[m]] = wg[m]
    L, _, _ = oracle.pbc_from_box(ls.box)
    eb = oracle.bonds(L, ls.bond_ids, ls.bond_k, ls.bond_r0, crd, w)
    return w, en, eb, float(en_abs[1]), npair


@pytest.mark.parametrize("gap", [1, 0])
def test_engine_hpsdh_forces_vs_oracle(gap):
    d = load_npz("wons_hpsdh_grad.npz")
    ls = cls_from_fixture(d)
    cfg = _cfg(d, **{"gap-bug": gap})
    say("HPS + Debye-Hueckel forces, gap rule", gap, ls.n_atoms)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    s = System(ls, cfg)
    w, e = s.energy_force(rebuild=True)
    s.close()
    wo, eo, eb, S, npair = _oracle_celllist(ls, crd, cfg, hps_dh_form(ls, d), bool(gap))
    err = rms_rel(w, wo)
    say("  pairs gpu %d oracle %d; gradient rms rel %.3e; vdw %.8f / %.8f; ele %.8f / %.8f (scale %.3e)" %
        (e["npairs"], npair, err, e["vdw"], eo[0], e["ele"], eo[1], S))
    assert abs(int(e["npairs"]) - npair) <= 1  # r^2 at the cutoff within an ulp may fall either side
    assert err < 2e-6
    assert e["vdw"] == pytest.approx(eo[0], rel=2e-6, abs=1e-6)
    assert abs(e["ele"] - eo[1]) <= 2e-6 * S
    assert e["bond"] == pytest.approx(eb, rel=1e-9, abs=1e-12)
    # no excluded-pair or self terms under Debye-Hueckel (ZeroMultipoleSum.cpp:75-82,456-466)
    assert e["excess"] == 0.0 and e["self"] == 0.0
    if gap == 0:
        # every pair within the cutoff: the reference all-pairs build's own gradient and log
        ref = d["work"]
        log0 = d["log"][0]
        say("  vs the reference run: gradient rms rel %.3e, vdw %.8f / %.8f, ele %.8f / %.8f" %
            (rms_rel(w, ref), e["vdw"], log0[4], e["ele"], log0[5]))
        assert rms_rel(w, ref) < 2e-6
        assert e["vdw"] == pytest.approx(log0[4], rel=2e-6, abs=1e-6)
        assert abs(e["ele"] - log0[5]) <= 2e-6 * S


def test_engine_hpsdh_trajectory_vs_reference():
    """20 leapfrog steps (dt 5 fs, the cg_q8 sample's ionic strength) against
    the reference all-pairs run: per-step potential and kinetic energy and the
    final restart coordinates."""
    d = load_npz("wons_hpsdh_traj.npz")
    ls = cls_from_fixture(d)
    cfg = _cfg(d, **{"gap-bug": 0})
    say("HPS + Debye-Hueckel trajectory", ls.n_atoms, cfg["time-step"])
    s = System(ls, cfg)
    s.run(20)
    log = s.energy_log()
    crd, _, _ = s.state()
    s.close()
    ref = d["log"]
    L, lb, ub = oracle.pbc_from_box(ls.box)
    _, _, _, S, _ = _oracle_celllist(ls, wrap_inbox(ls.crd, L, lb, ub), cfg, hps_dh_form(ls, d), False)
    pot = log[:, 0] + log[:, 1] + log[:, 2] + log[:, 3] + log[:, 4]
    for k in range(20):
        tol = 2e-6 * abs(ref[k, 4]) + 2e-6 * S + 1e-9 * abs(ref[k, 3]) + 1e-6
        if k in (0, 19):
            say("  step %d pot %.8f ref %.8f (tol %.2e) kin %.8f ref %.8f" % (k, pot[k], ref[k, 1], tol, log[k, 5],
                                                                           ref[k, 2]))
        assert abs(pot[k] - ref[k, 1]) <= tol, k
        assert log[k, 5] == pytest.approx(ref[k, 2], rel=1e-6), k
    say("  max |dx| vs reference %.3e" % np.max(np.abs(crd - d["crd"])))
    assert np.max(np.abs(crd - d["crd"])) < 1e-5


def test_engine_dh_gradient_fix_vs_oracle():
    """dh-grad-fix 1: the Debye-Hueckel gradient coefficient -E (1/r + kappa) / r
    (the derivative of E) in place of the reference's -E (1/r + kappa)."""
    d = load_npz("wons_hpsdh_grad.npz")
    ls = cls_from_fixture(d)
    cfg = _cfg(d, **{"gap-bug": 0, "dh-grad-fix": 1})
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    s = System(ls, cfg)
    w, e = s.energy_force(rebuild=True)
    s.close()
    wo, eo, _, S, _ = _oracle_celllist(ls, crd, cfg, hps_dh_form(ls, d, dh_grad_fix=True), False)
    w_ref, _, _, _, _ = _oracle_celllist(ls, crd, cfg, hps_dh_form(ls, d), False)
    say("  gradient fix: rms rel vs oracle %.3e (the reference gradient differs by %.3e)" %
        (rms_rel(w, wo), rms_rel(w_ref, wo)))
    assert rms_rel(w, wo) < 2e-6
    assert rms_rel(w_ref, wo) > 1e-3
    assert abs(e["ele"] - eo[1]) <= 2e-6 * S


def test_call_points_hpsdh_vs_oracle():
    """The reference's own call sequence for the HPS + Debye-Hueckel build
    (og_alloc_set_hps_params, og_hps_constant, og_debye_huckel_constant,
    og_pairwise_hps_dh, og_memcpy_dtoh_work) on the host-built grid, against
    the oracle's cell-list loop on the reference list (gap rule on)."""
    d = load_npz("wons_hpsdh_grad.npz")
    ls = cls_from_fixture(d)
    cfg = _cfg(d)
    n = ls.n_atoms
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    zd = oracle.zd_setup(cfg["cutoff"], 0.0)
    o = oracle.Oracle(n, ls.box, cfg["nsgrid-cutoff"])
    c = o.build(crd)
    g = o.g
    cell_col = c.cell_crd[:, 0] + c.cell_crd[:, 1] * g.n_cells_xyz[0]
    grid = {"n_cells": c.n_cells, "atomids": c.atomids, "crd": c.crd, "idx_xy_head_cell": c.idx_xy_head_cell,
            "column_of_slot": np.repeat(cell_col, 8).astype(np.int32)}
    geom = {"n_cells_xyz": list(g.n_cells_xyz), "L_cell_xyz": list(g.L_cell_xyz),
            "n_neighbors_xyz": list(g.n_neighbors_xyz), "n_columns": g.n_columns, "max_n_cells": g.max_n_cells,
            "max_n_cell_pairs": g.max_n_cell_pairs}
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    hc, hl = ls.hps_tables_f32()
    drv = SubBoxDriver(n, L, lb, cfg["cutoff"], cfg["nsgrid-cutoff"], zd, lj6, lj12, ls.nb15off_table(), q,
                       ls.atom_type, geom, max_n_atom_array=n + 8 * g.n_columns)
    try:
        drv.set_hps_dh(hc, hl, cfg["hydrophobicity-scale-epsiron"], cfg["debye-huckel-dielectric"],
                       cfg["debye-huckel-temperature"], cfg["debye-huckel-ionic-strength"])
        drv.update_grid(grid)
        w, e = drv.pair_forces()
    finally:
        drv.close()
    cps, _ = o.enumerate(ls.nb15off_table(), gap_bug=True)
    ff = hps_dh_form(ls, d)
    wo, eo, _ = o.pairs_ff(zd, ff, cps, ls.atom_type, lj6, lj12, q)
    _, eabs, _ = o.pairs_ff(zd, ff, cps, ls.atom_type, lj6, lj12, np.abs(q))
    m = c.atomids >= 0
    say("  call points: rms rel %.3e, vdw %.8f / %.8f, ele %.8f / %.8f" % (rms_rel(w[m], wo[m]), e[0], eo[0], e[1],
                                                                        eo[1]))
    assert rms_rel(w[m], wo[m]) < 2e-6
    assert np.all(w[~m] == 0.0)
    assert e[0] == pytest.approx(eo[0], rel=2e-6, abs=1e-6)
    assert abs(e[1] - eo[1]) <= 2e-6 * eabs[1]
