"""Angles, torsions, impropers, scaled 1-4 pairs and four-atom SHAKE on the
device (og_bonded.hip; the reference's host terms SubBox.cpp:1250-1346 with
ForceField.cpp:138-331, ConstraintShake::shake_quad ConstraintShake.cpp:151-276)
against the oracle and against the reference all-pairs build's own runs of
the ala3 launch set (tests/golden/wons_ala3_*.npz).

Bars: the bonded gradient (a run with the terms less a run without them;
the pair part is the same fixed-point sum in both) to 1e-12 RMS relative of
the oracle's FP64 gradient; bonded energies to 1e-10 relative (+1e-8 for the
2^-32 fixed-point energy sums); pair terms as the pair-kernel tests (FP32)."""
import copy
import sys

import numpy as np
import pytest

import oracle
from helpers import cls_from_fixture, ele_scale, load_npz, rms_rel, wrap_inbox
from omegagene_amd.engine import Batch, System, potential

pytestmark = pytest.mark.gpu

BONDED = ("angle", "torsion", "impro", "vdw14", "ele14")
CFG = {"cutoff": 12.0, "nsgrid-cutoff": 13.0, "nsgrid-update-intvl": 1, "time-step": 2.0, "ele-alpha": 0.0,
       "integrator": "leapfrog-presto", "const-tolerance": 1e-10, "const-max-loops": 1000, "gap-bug": 0,
       "log-cap": 64}


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _without_bonded(ls):
    x = copy.copy(ls)
    x.angles, x.torsions, x.impros, x.nb14 = [], [], [], []
    return x


def _raw_log_terms(d):
    """Angle, Torsion, Improper, 14-VDW, 14-Ele per step from the reference log lines."""
    out = []
    for ln in bytes(d["raw_log"]).decode().splitlines():
        f = ln.split()
        if ln.startswith("Bond:"):
            out.append({"bond": float(f[1]), "angle": float(f[3])})
        elif ln.startswith("Torsion:"):
            out[-1].update(torsion=float(f[1]), impro=float(f[3]))
        elif ln.startswith("14-VDW:"):
            out[-1].update(vdw14=float(f[1]), ele14=float(f[3]))
    return out


def test_ala3_bonded_gradient_and_energies_vs_oracle():
    d = load_npz("wons_ala3_grad.npz")
    ls = cls_from_fixture(d)
    assert len(ls.angles) and len(ls.torsions) and len(ls.impros) and len(ls.nb14)
    cfg = dict(CFG, constraint="none")
    say("ala3 bonded terms", ls.n_atoms, len(ls.angles), len(ls.torsions), len(ls.impros), len(ls.nb14))
    s = System(ls, cfg)
    w, e = s.energy_force(rebuild=True)
    s.close()
    s0 = System(_without_bonded(ls), cfg)
    w0, e0 = s0.energy_force(rebuild=True)
    s0.close()
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    wb = np.zeros((ls.n_atoms, 3))
    eo = oracle.bonded(ls, crd, wb)
    err = rms_rel(w - w0, wb)
    say("  bonded gradient rms rel %.3e; energies %s" %
        (err, ", ".join("%s %.10f/%.10f" % (k, e[k], eo[k]) for k in BONDED)))
    assert err < 1e-12
    for k in BONDED:
        assert e[k] == pytest.approx(eo[k], rel=1e-10, abs=1e-8), k
        assert e0[k] == 0.0
    # the pair, excluded-pair and bond terms are the same with and without the bonded terms
    for k in ("vdw", "ele", "excess", "self", "bond", "npairs"):
        assert e[k] == e0[k], k
    # the whole gradient against the reference run's (pairs in FP32 there too: the pair-kernel bar)
    say("  whole gradient vs the reference run: rms rel %.3e" % rms_rel(w, d["work"]))
    assert rms_rel(w, d["work"]) < 2e-6
    log0 = d["log"][0]  # step, pot, kin, bond, vdw, ele
    tol = 2e-6 * abs(log0[4]) + 1e-8 * ele_scale(ls, crd=crd) + 1e-8
    say("  potential %.8f reference %.8f (tol %.2e)" % (potential(e), log0[1], tol))
    assert abs(potential(e) - log0[1]) <= tol


def test_ala3_md_with_four_atom_shake_vs_reference_run():
    """10 leapfrog steps with SHAKE pairs, trios and four-atom groups and COM
    cancel on the solute and solvent groups: per-step potential, kinetic and
    bonded energies against the reference run's log, the final restart
    coordinates, and every four-atom group at its six target distances."""
    d = load_npz("wons_ala3_traj.npz")
    ls = cls_from_fixture(d)
    assert len(ls.shake_quads) == 9
    groups = [g for name, g in ls.groups if name in ("solute", "solvent")]
    cfg = dict(CFG, constraint="shake", **{"com-motion": "cancel"})
    say("ala3 MD with four-atom SHAKE", ls.n_atoms, len(ls.shake_pair_ids), len(ls.shake_trio_ids),
        len(ls.shake_quads))
    s = System(ls, cfg, com_groups=groups)
    s.run(10)
    log = s.energy_log()
    crd, vn, _ = s.state()
    s.close()
    ref = d["log"]
    raw = _raw_log_terms(d)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    S = ele_scale(ls, crd=wrap_inbox(ls.crd, L, lb, ub))
    pot = potential(log)
    cols = {k: i for i, k in enumerate(["vdw", "ele", "excess", "self", "bond", "kinetic", "virial", "extend",
                                        "npairs", "lambda", "vstate", "angle", "torsion", "impro", "vdw14",
                                        "ele14"])}
    for k in range(10):
        tol = 2e-6 * abs(ref[k, 4]) + 1e-8 * S + 1e-6
        if k in (0, 9):
            say("  step %d pot %.8f ref %.8f (tol %.2e) kin %.8f ref %.8f" % (k, pot[k], ref[k, 1], tol, log[k, 5],
                                                                           ref[k, 2]))
        assert abs(pot[k] - ref[k, 1]) <= tol, k
        assert log[k, 5] == pytest.approx(ref[k, 2], rel=1e-6), k
        for t in BONDED:
            # the log prints 11 significant digits; later steps carry the FP32 pair differences of the trajectory
            bar = (1e-9 if k == 0 else 1e-6) * abs(raw[k][t]) + 1e-8
            assert abs(log[k, cols[t]] - raw[k][t]) <= bar, (k, t, log[k, cols[t]], raw[k][t])
    dx = np.max(np.abs(crd - d["crd"]))
    say("  max |dx| vs reference %.3e" % dx)
    assert dx < 1e-5
    ids, d2 = ls.shake_quad_tables()
    pairs = [(0, 1), (1, 2), (2, 0), (3, 0), (2, 3), (1, 3)]
    Lv = np.asarray(L, np.float64)
    worst = 0.0
    for q in range(len(ids)):
        for p, (i, j) in enumerate(pairs):
            v = crd[ids[q, i]] - crd[ids[q, j]]
            v -= Lv * np.round(v / Lv)
            worst = max(worst, abs(v @ v - d2[q, p]) / d2[q, p])
    say("  four-atom groups: worst relative squared-distance error %.3e" % worst)
    assert worst < 1e-9


def test_ala3_batch_equals_single_bitwise():
    """Two copies of the bonded system (one with other velocities) in one batch
    against each run alone: the same bits."""
    d = load_npz("wons_ala3_traj.npz")
    ls = cls_from_fixture(d)
    ls2 = copy.copy(ls)
    ls2.vel = ls.vel * 0.5
    groups = [g for name, g in ls.groups if name in ("solute", "solvent")]
    cfg = dict(CFG, constraint="shake", **{"com-motion": "cancel"})
    alone = []
    for x in (ls, ls2):
        s = System(x, cfg, com_groups=groups)
        s.run(5)
        alone.append((s.state()[0], s.energy_log()))
        s.close()
    b = Batch()
    ss = [System(x, cfg, com_groups=groups) for x in (ls, ls2)]
    for s in ss:
        b.add(s)
    b.run(5)
    for s, (c, lg) in zip(ss, alone):
        assert np.array_equal(s.state()[0], c)
        assert np.array_equal(s.energy_log(), lg)
    b.close()
    for s in ss:
        s.close()
