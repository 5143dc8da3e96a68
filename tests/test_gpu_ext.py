"""V-McMD and V-AUS on the device (k_vmcmd / k_vmcmd_scale, og_vaus_lane0;
the reference's ExtendedVMcMD / ExtendedVAUS, Extend.cpp:126-540) against
the reference all-pairs build's own runs (tests/golden/wons_ext.npz) and the
fused Langevin step against the separate kernels.

Bars: lambda per step against the reference's lambda file within the FP32
pair-sum scale (V-McMD: 2e-6 of the vdw energy + 1e-8 of the ele term
magnitudes; V-AUS: 1e-6 A); the virtual state of every step equal to the
reference's (the fixtures' transition probabilities are 0 or 1, so the walk
does not depend on the random stream); per-step potential and kinetic energy
and the final coordinates within 2e-5 A."""
import sys

import numpy as np
import pytest

import oracle
from helpers import ele_scale, load_npz
from omegagene_amd import gen
from omegagene_amd.cls import GroupCoordDef, VmcmdDef, read_cls
from omegagene_amd.engine import Batch, System, potential

pytestmark = pytest.mark.gpu

BASE = {"cutoff": 12.0, "nsgrid-cutoff": 13.0, "nsgrid-update-intvl": 1, "time-step": 2.0, "gap-bug": 0,
        "integrator": "leapfrog-presto", "constraint": "shake", "const-tolerance": 1e-10, "temperature": 300.0,
        "log-cap": 64}
LAMBDA, VSTATE = 9, 10   # energy-log columns


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("tag", ["mcmd", "aus", "aus_rec"])
def test_vmcmd_vaus_vs_reference_run(tag):
    d = load_npz("wons_ext.npz")
    ls = read_cls(bytes(d[tag + "_cls"]))
    mode = "v-mcmd" if tag == "mcmd" else "v-aus"
    cfg = dict(BASE, **{"extended-ensemble": mode})
    if mode == "v-aus":
        cfg.update({"enhance-sigma": 0.5, "enhance-recovery-coef": 2.0})
    say(mode, tag, ls.n_atoms, len(ls.vmcmd.states), "states")
    s = System(ls, cfg)
    s.run(20)
    log = s.energy_log()
    crd, _, _ = s.state()
    s.close()
    ref = d[tag + "_log"]
    lam_ref = d[tag + "_lambda"]
    S = ele_scale(ls)
    pot = potential(log)
    vs_ref = {int(st): int(v) - 1 for st, v in d[tag + "_vs"]}
    state = ls.vmcmd.init_vs
    worst = 0.0
    for k in range(20):
        tol = 2e-6 * abs(ref[k, 4]) + 1e-8 * S + 1e-6
        assert abs(pot[k] - ref[k, 1]) <= tol, k
        assert log[k, 5] == pytest.approx(ref[k, 2], rel=1e-6), k
        state = vs_ref.get(k, state)
        assert int(log[k, VSTATE]) == state, (k, log[k, VSTATE], state)
        if k >= 1:
            lt = tol if mode == "v-mcmd" else 1e-6
            worst = max(worst, abs(log[k, LAMBDA] - lam_ref[k - 1]) / lt)
            assert abs(log[k, LAMBDA] - lam_ref[k - 1]) <= lt, (k, log[k, LAMBDA], lam_ref[k - 1])
    dx = np.max(np.abs(crd - d[tag + "_crd"]))
    say("  states %s; lambda worst |diff|/tol %.2e; max |dx| vs reference %.3e" %
        (sorted(set(int(v) for v in log[:, VSTATE])), worst, dx))
    # 20 steps from a random water lattice (forces of the first steps are large) with the gradient scaled by
    # a factor of the FP32-summed potential: 1.0e-5 A measured, 2e-5 allowed
    assert dx < 2e-5
    assert log[:, 7].max() == 0.0   # the reference reports no bias energy for V-McMD / V-AUS


def _vaus_walker(n, seed, cfg):
    """A walker with V-AUS on its chain's end groups: three states around the
    starting distance (the device's own lambda of step 0, minimum image
    included), moves likely both ways."""
    ls = gen.walker_system(n, seed=seed, vcmd=True)
    ls.vcmd = None
    names = [nm for nm, _ in ls.groups]
    ga, gb = dict(ls.groups)["end_a"], dict(ls.groups)["end_b"]
    ls.group_coord = GroupCoordDef("V-AUS restart", 0, [names.index("end_a") + 1, names.index("end_b") + 1],
                                   [ls.crd[ga].copy(), ls.crd[gb].copy()])
    ls.vmcmd = VmcmdDef(1000, 300.0, [(0, 0.0, 1000.0, 0.0, 0.0, [0.0], 0.0, 0.0)], 0, 5)
    s = System(ls, dict(cfg, **{"constraint": "none"}))
    s.run(1)
    lam0 = float(s.energy_log()[0, LAMBDA])
    s.close()
    cuts = [lam0 - 4.0, lam0 - 0.5, lam0 + 0.5, lam0 + 4.0]
    states = [(2, cuts[k] - 0.6, cuts[k + 1] + 0.6, 0.5, 0.5, [0.3 - 0.2 * k, -0.02, 0.001], 0.0, 0.0)
              for k in range(3)]
    ls.vmcmd = VmcmdDef(4, 300.0, states, 1, 5)
    ls.settle_ids, ls.settle_d = ls.shake_trio_ids, ls.shake_trio_d
    ls.shake_trio_ids = np.zeros((0, 3), np.int32)
    ls.shake_trio_d = np.zeros((0, 3))
    return ls


def test_vaus_fused_langevin_step_equals_split_kernels():
    """V-AUS inside the one-kernel Langevin step (its free-atom block) gives
    bitwise the trajectory, lambda and state walk of the separate kernels,
    alone and in a batch; the walk moves (probabilities 0.5)."""
    cfg = dict(BASE, **{"integrator": "langevin", "constraint": "shake-settle", "langevin-gamma": 20.0,
                        "nsgrid-update-intvl": 5, "random-seed": 3, "extended-ensemble": "v-aus",
                        "enhance-sigma": 0.5, "enhance-recovery-coef": 1.0})
    lss = [_vaus_walker(n, 91 + k, cfg) for k, n in enumerate((4000, 6500))]
    runs = {}
    for split in (1, 0):
        out = []
        for ls in lss:
            s = System(ls, dict(cfg, **{"split-step": split}))
            assert s.step_fused() == (split == 0)
            s.run(40)
            out.append((s.state(), s.energy_log()))
            s.close()
        runs[split] = out
    b = Batch()
    ss = [System(ls, cfg) for ls in lss]
    for s in ss:
        b.add(s)
    b.run(40)
    batch = [(s.state(), s.energy_log()) for s in ss]
    b.close()
    for s in ss:
        s.close()
    grouped = [2, 4, 5]  # excess, bond, kinetic: block sums grouped by water in the fused step
    for (st1, log1), (st0, log0), (stb, logb) in zip(runs[1], runs[0], batch):
        for a, c, e in zip(st1, st0, stb):
            assert np.array_equal(a, c) and np.array_equal(c, e)
        other = [k for k in range(log1.shape[1]) if k not in grouped]
        assert np.array_equal(log1[:, other], log0[:, other])
        assert np.array_equal(log0, logb)
        say("  V-AUS fused = split over 40 steps; states visited %s, lambda %.4f .. %.4f" %
            (sorted(set(int(v) for v in log0[:, VSTATE])), log0[:, LAMBDA].min(), log0[:, LAMBDA].max()))
    assert any(len(set(int(v) for v in r[1][:, VSTATE])) > 1 for r in runs[0])


def test_vmcmd_factor_scales_the_first_step():
    """V-McMD over one leapfrog step without constraints: the velocity change
    is the unbiased one times k_B T d ln P(lambda) of the initial state at the
    device's own lambda (the factor of Extend.cpp:177-200 on the oracle)."""
    d = load_npz("wons_ext.npz")
    ls = read_cls(bytes(d["mcmd_cls"]))
    cfg = dict(BASE, **{"constraint": "none", "com-motion": "none"})
    s0 = System(ls, cfg)
    s0.run(1)
    _, v_plain, _ = s0.state()
    s0.close()
    s1 = System(ls, dict(cfg, **{"extended-ensemble": "v-mcmd"}))
    s1.run(1)
    _, v_bias, _ = s1.state()
    row = s1.energy_log()[0]
    s1.close()
    st = ls.vmcmd.states[ls.vmcmd.init_vs]
    f = oracle.vmcmd_factor(st[5], st[1], st[2], ls.vmcmd.temperature, row[LAMBDA])
    dv_plain = v_plain - ls.vel
    dv_bias = v_bias - ls.vel
    err = np.max(np.abs(dv_bias - f * dv_plain)) / np.max(np.abs(dv_plain))
    say("  V-McMD factor %.9f at lambda %.6f (state %d); max |dv - f dv0| / max |dv0| = %.2e" %
        (f, row[LAMBDA], int(row[VSTATE]), err))
    assert int(row[VSTATE]) == ls.vmcmd.init_vs
    assert abs(f - 1.0) > 1e-3
    assert err < 1e-12
