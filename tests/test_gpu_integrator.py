"""Device integrator pieces against the oracle on identical inputs, and the
reference's own step-loop runs (tests/golden/wons_*.npz) through the HIP path.

One-step comparisons feed the oracle the device's own gradient of the step
(og_get_work), so the FP64 integrator, constraint, COM, thermostat and
Langevin arithmetic is compared bit for bit (<= 1e-12 A; the COM and
thermostat sums are fixed-order trees on the device and atom-order loops in
the reference, so those two are compared to 1e-12 relative).  Multi-step
runs against the reference differ only by the FP32 pair sums (order), so they
are compared like the existing 20-step reference trajectory."""
import sys

import numpy as np
import pytest

import oracle
from helpers import ala3_launchset, chain_shake_pairs, cls_from_fixture, ele_scale, load_npz, wrap_inbox
from omegagene_amd import gen
from omegagene_amd.engine import System

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


BASE = {"cutoff": 12.0, "nsgrid-cutoff": 13.0, "nsgrid-update-intvl": 1, "time-step": 2.0,
        "integrator": "leapfrog-presto", "constraint": "none", "const-tolerance": 1e-10, "ele-alpha": 0.0,
        "gap-bug": 0}


def _settle(ls):
    ls.settle_ids, ls.settle_d = ls.shake_trio_ids, ls.shake_trio_d
    ls.shake_trio_ids = np.zeros((0, 3), np.int32)
    ls.shake_trio_d = np.zeros((0, 3))
    return ls


def _masses(ls):
    m32 = ls.mass.astype(np.float32)
    return m32, (1.0 / m32.astype(np.float64)).astype(np.float32)


@pytest.mark.parametrize("constraint", ["none", "shake-settle"])
def test_langevin_with_reference_noise_bitwise_vs_oracle(constraint):
    """Position Langevin at gamma 1/ps fed the reference's noise stream
    (og_set_langevin_noise with orc_ref_normals), 4 steps including the
    first-step Euler: positions and central-difference velocities equal the
    oracle's (SubBox.cpp:2054-2114) given the device gradient.  Without
    constraints the separate kernels run; with SETTLE waters the fused step
    (SETTLE after the Langevin update, a documented extension: the reference
    Langevin step applies no constraint, DynamicsMode.cpp:877)."""
    ls = gen.build_system(400, seed=71, n_beads=20)
    if constraint != "none":
        ls = _settle(ls)
    gamma, T, seed = 1.0, 300.0, 4321
    cfg = dict(BASE, **{"integrator": "langevin", "constraint": constraint, "langevin-gamma": gamma,
                        "temperature": T, "random-seed": 99})
    s = System(ls, cfg)
    say("langevin injected noise", constraint, ls.n_atoms, "fused" if s.step_fused() else "split")
    assert s.step_fused() == (constraint != "none")
    L, lb, ub = oracle.pbc_from_box(ls.box)
    m32, minv = _masses(ls)
    n = ls.n_atoms
    x = wrap_inbox(ls.crd, L, lb, ub)
    xp = x.copy()
    for step in range(4):
        if step > 0:
            zeta = oracle.ref_normals(seed, (step - 1) * 3 * n, 3 * n).reshape(n, 3)
            s.set_langevin_noise(zeta)
        s.run(1)
        w = s.work()
        crd_d, _, vj_d = s.state()
        xp2, xp = xp, x
        if step == 0:
            xn = oracle.update_coordinates(2.0, xp, ls.vel)
        else:
            xn = oracle.langevin_positions(1.0, gamma / 1000.0, T, L, xp, xp2, w, m32, minv, zeta)
        if constraint != "none":
            xn = oracle.settle(L, ls.settle_ids, ls.settle_d[0][0], ls.settle_d[0][1], xn, xp, m32)
        if step > 0:
            v = oracle.langevin_velocity(2.0, L, xn, xp2)
            dv = np.max(np.abs(v - vj_d))
        else:
            dv = np.max(np.abs(ls.vel - vj_d))
        x = oracle.revise_pbc(L, lb, xn)
        dx = np.max(np.abs(x - crd_d))
        say("  step %d max|dx| %.2e max|dv| %.2e" % (step, dx, dv))
        assert dx <= 1e-12 and dv <= 1e-14
    # the injected noise acts: without it the next step lands elsewhere
    s.close()


def test_langevin_reference_run_through_device():
    """tests/golden/wons_langevin.npz (reference, gamma 10/ps, seed 1234, no
    constraints, 3 steps + the extra last step) through the device with the
    reference noise injected: the restart's x_prev and v agree to the FP32
    pair-sum scale, the logged potential per step to the usual bars."""
    d = load_npz("wons_langevin.npz")
    ls = cls_from_fixture(d)
    n = ls.n_atoms
    cfg = dict(BASE, **{"integrator": "langevin", "langevin-gamma": 10.0, "temperature": 300.0, "random-seed": 5,
                        "log-cap": 8, "gap-bug": 0})
    s = System(ls, cfg)
    say("langevin reference run", n)
    x3 = None
    for step in range(4):
        if step > 0:
            s.set_langevin_noise(oracle.ref_normals(1234, (step - 1) * 3 * n, 3 * n))
        s.run(1)
        if step == 2:
            x3 = s.state()[0]
    _, _, v3 = s.state()
    log = s.energy_log()
    S = ele_scale(ls)
    for k in range(4):
        pot = log[k, 0] + log[k, 1] + log[k, 2] + log[k, 3] + log[k, 4]
        assert abs(pot - d["log"][k, 1]) <= 2e-6 * abs(d["log"][k, 4]) + 1e-8 * S, k
    dx, dv = np.max(np.abs(x3 - d["crd"])), np.max(np.abs(v3 - d["vel"]))
    say("  max|dx| %.2e max|dv| %.2e vs reference restart" % (dx, dv))
    assert dx < 1e-5 and dv < 1e-5
    s.close()


def _one_leapfrog_step(ls, cfg, com_groups=None, thermo=None):
    """Device step vs the oracle's step (DynamicsModePresto order) given the device gradient."""
    s = System(ls, cfg)
    crd0, vn0, _ = s.state()
    s.run(1)
    w = s.work()
    crd1, vn1, vj1 = s.state()
    L, lb, ub = oracle.pbc_from_box(ls.box)
    m32, minv = _masses(ls)
    v = oracle.update_velocities(2.0, vn0, w, minv)
    if com_groups:
        v = oracle.cancel_com(com_groups, m32, v)
    if thermo:
        T = thermo["temperature_init"]  # step 0 of the heating ramp
        v, _ = oracle.thermostat_scaling(2.0, oracle.thermostat_tau_inv(2.0, thermo["tau_fs"]), T,
                                         oracle.temperature_coeff(ls.d_free(False)), w, vn0, m32, minv)
    x = oracle.update_coordinates(2.0, crd0, v)
    if cfg["constraint"] != "none":
        if len(ls.shake_pair_ids):
            x = oracle.shake_pair(L, ls.shake_pair_ids, np.asarray(ls.shake_pair_d) ** 2, x, crd0, minv)
        if len(ls.shake_trio_ids):
            x, nf = oracle.shake_trio(L, ls.shake_trio_ids, np.asarray(ls.shake_trio_d) ** 2, x, crd0, minv,
                                      1e-10, 1000)
            assert nf == 0
        if len(ls.settle_ids):
            x = oracle.settle(L, ls.settle_ids, ls.settle_d[0][0], ls.settle_d[0][1], x, crd0, m32)
        v = oracle.velocity_from_crd(2.0, x, crd0)
    x = oracle.revise_pbc(L, lb, x)
    s.close()
    return x, v, crd1, vn1


def test_shake_pair_step_vs_oracle():
    """SHAKE pairs on a chain (disjoint, ConstraintShake.cpp:18-49) and trios on
    the waters: one device step equals the oracle's given the same gradient."""
    ls = chain_shake_pairs(gen.build_system(500, seed=43, n_beads=40))
    x, v, x1, v1 = _one_leapfrog_step(ls, dict(BASE, constraint="shake"))
    dx, dv = np.max(np.abs(x - x1)), np.max(np.abs(v - v1))
    say("shake pair step max|dx| %.2e max|dv| %.2e" % (dx, dv))
    assert dx <= 1e-12 and dv <= 1e-12


def test_settle_step_vs_oracle():
    """Analytic SETTLE on the device (og_settle_core) equals orc_settle given
    the same unconstrained positions (one leapfrog step, separate kernels)."""
    ls = _settle(gen.build_system(500, seed=44))
    x, v, x1, v1 = _one_leapfrog_step(ls, dict(BASE, constraint="shake-settle"))
    dx, dv = np.max(np.abs(x - x1)), np.max(np.abs(v - v1))
    say("settle step max|dx| %.2e max|dv| %.2e" % (dx, dv))
    assert dx <= 1e-12 and dv <= 1e-12


def _water_across_the_face(ls, dt=2.0, gap=1e-8):
    """Water 0 of ls made non-interacting (zero charges, a new LJ type with
    zero coefficients, so its gradient is exactly zero) and placed rigid with
    its O 0.01 A below the top y face and H1 0.02 A above the bottom one (0.03
    A apart in y across the face); H1 moves down at a speed that puts its
    unconstrained position gap - L from O's in y after one leapfrog step --
    within float rounding of -L, where the reference's two-pass minimum image
    returns a vector one box long."""
    L = float(ls.box[4])
    th = np.radians(104.52)
    dOH = gen.D_OH
    o, h1, h2 = ls.settle_ids[0]
    O = np.array([0.4 * L, L - 0.01, 0.6 * L])
    u1 = np.array([np.sqrt(dOH ** 2 - 0.03 ** 2), 0.03, 0.0]) / dOH
    w = np.array([O, O + dOH * u1, O + dOH * (np.cos(th) * u1 + np.sin(th) * np.array([0.0, 0.0, 1.0]))])
    w[1, 1] -= L
    ls.crd[[o, h1, h2]] = w
    ls.vel[[o, h1, h2]] = 0.0
    ls.vel[h1, 1] = -(0.03 - gap) / dt
    nt = ls.n_lj_types
    ls.lj6 = np.pad(ls.lj6, ((0, 1), (0, 1)))
    ls.lj12 = np.pad(ls.lj12, ((0, 1), (0, 1)))
    ls.n_lj_types = nt + 1
    ls.atom_type[[o, h1, h2]] = nt
    ls.charge[[o, h1, h2]] = 0.0
    return (o, h1, h2)


def _rigid_error(x, ids, L):
    d = lambda a, b: np.linalg.norm((x[a] - x[b]) - L * np.round((x[a] - x[b]) / L))
    o, h1, h2 = ids
    return max(abs(d(o, h1) - gen.D_OH), abs(d(o, h2) - gen.D_OH), abs(d(h1, h2) - gen.D_HH))


def test_settle_across_the_box_face():
    """A water whose unconstrained H1 - O difference in y lands within float
    rounding of -L (H1 just below the bottom face, O just below the top one):
    the device SETTLE keeps it rigid with no constraint failure, and the step
    equals the oracle's (both take one nearest-image step more where the
    reference's two-pass minimum image returns a vector one box long)."""
    ls = _settle(gen.build_system(600, seed=48, L=30.0))
    ids = _water_across_the_face(ls)
    L = float(ls.box[0])
    x, v, x1, v1 = _one_leapfrog_step(ls, dict(BASE, constraint="shake-settle"))
    dx, dv = np.max(np.abs(x - x1)), np.max(np.abs(v - v1))
    say("water across the y face: leapfrog step vs oracle max|dx| %.2e max|dv| %.2e; rigid error %.2e; "
        "O y %.6f H1 y %.6f" % (dx, dv, _rigid_error(x1, ids, L), x1[ids[0], 1], x1[ids[1], 1]))
    assert dx <= 1e-12 and dv <= 1e-12
    assert _rigid_error(x1, ids, L) < 1e-10


def test_com_cancel_step_vs_oracle():
    """--com-motion cancel on two groups (COMMotion.cpp:30-55) with SHAKE trios:
    one device step against the oracle; each group's momentum is zero after."""
    ls = gen.build_system(500, seed=45, n_beads=20)
    groups = [g for _, g in ls.groups]
    x, v, x1, v1 = _one_leapfrog_step(ls, dict(BASE, **{"constraint": "shake", "com-motion": "cancel"}),
                                      com_groups=groups)
    dx = np.max(np.abs(x - x1))
    dv = np.max(np.abs(v - v1)) / np.max(np.abs(v))
    say("com cancel step max|dx| %.2e rel max|dv| %.2e" % (dx, dv))
    assert dx <= 1e-12 and dv <= 1e-12


def test_thermostat_step_vs_oracle():
    """--thermostat scaling (Thermostat.cpp:57-84), no constraints: one device
    step against the oracle, at the first temperature of a heating ramp."""
    ls = gen.build_system(500, seed=47, n_beads=20)
    th = dict(tau_fs=100.0, temperature_init=250.0)
    cfg = dict(BASE, **{"thermostat": "scaling", "berendsen-tau": 0.1, "temperature": 300.0,
                        "temperature-init": 250.0, "heating-steps": 5})
    x, v, x1, v1 = _one_leapfrog_step(ls, cfg, thermo=th)
    dx = np.max(np.abs(x - x1))
    dv = np.max(np.abs(v - v1)) / np.max(np.abs(v))
    say("thermostat step max|dx| %.2e rel max|dv| %.2e" % (dx, dv))
    assert dx <= 1e-12 and dv <= 1e-12


@pytest.mark.parametrize("name,extra", [
    ("wons_shake_pair", {"constraint": "shake"}),
    ("wons_com", {"constraint": "shake", "com-motion": "cancel"}),
    ("wons_thermo", {"thermostat": "scaling", "berendsen-tau": 0.1, "temperature": 300.0,
                     "temperature-init": 250.0, "heating-steps": 5}),
])
def test_reference_runs_through_device(name, extra):
    """The reference's 10-step runs (SHAKE pairs, COM cancel, scaling
    thermostat) through the device: per-step potential to the FP32 pair-sum
    bars, kinetic energy to 1e-6, final restart within 1e-5 A."""
    d = load_npz(name + ".npz")
    ls = cls_from_fixture(d)
    s = System(ls, dict(BASE, **dict(extra, **{"log-cap": 16})))
    s.run(10)
    log = s.energy_log()
    S = ele_scale(ls)
    ref = d["log"]
    for k in range(10):
        pot = log[k, 0] + log[k, 1] + log[k, 2] + log[k, 3] + log[k, 4]
        assert abs(pot - ref[k, 1]) <= 2e-6 * abs(ref[k, 4]) + 1e-8 * S, (k, pot, ref[k, 1])
        assert log[k, 5] == pytest.approx(ref[k, 2], rel=1e-6), k
    crd, vn, _ = s.state()
    dx = np.max(np.abs(crd - d["crd"]))
    say("%s 10 steps: max|dx| vs reference %.2e" % (name, dx))
    assert dx < 1e-5
    s.close()


def test_ala3_pair_energies_through_device():
    """The reference's only real sample (samples/ala3, 6,207 atoms, 40 LJ types,
    its exclusion lists): one device force evaluation against the logged
    step-0 VDW 4164.8848106 and Ele -29051.201002 (log03_md.txt:133), which
    come from the all-pairs build (the full z-bin scan here, so the list
    holds every pair the all-pairs loop evaluates)."""
    ls, exp = ala3_launchset()
    s = System(ls, dict(BASE, **{"constraint": "none"}))
    w, e = s.energy_force(rebuild=True)
    s.close()
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    zd = oracle.zd_setup(12.0, 0.0)
    lj6, lj12 = ls.lj_tables_f32()
    _, en, npair = oracle.allpairs_fp32(zd, L, crd, ls.atom_type, lj6, lj12, ls.charge.astype(np.float32),
                                        ls.nb15off_table())
    ele = e["ele"] + e["excess"] + e["self"]
    S = ele_scale(ls)
    say("ala3: pairs %d (all-pairs %d), VDW %.7f (log %.7f), Ele %.6f (log %.6f)" %
        (e["npairs"], npair, e["vdw"], exp["vdw"], ele, exp["ele"]))
    assert abs(int(e["npairs"]) - int(npair)) <= 2
    assert e["vdw"] == pytest.approx(exp["vdw"], rel=2e-6)
    assert abs(ele - exp["ele"]) <= 1e-8 * S + 1e-6
