"""Batches and serving under the conditions the walker and serving
workloads meet: members at different rebuild phases, a member that faults,
new coordinates mid-run, resubmission under load, and the full 64-walker
batch in 500-step graphs.  Every comparison is bitwise: forces are
fixed-point integer sums and the Langevin noise is keyed by (seed, step,
atom), so a walker's trajectory does not depend on its company."""
import sys

import numpy as np
import pytest

from omegagene_amd import gen
from omegagene_amd.engine import OG_E_NAN, Batch, EngineError, Server, System

pytestmark = pytest.mark.gpu

CFG = {"cutoff": 12.0, "nsgrid-cutoff": 13.0, "nsgrid-update-intvl": 10, "time-step": 2.0, "integrator": "langevin",
       "constraint": "shake-settle", "const-tolerance": 1e-8, "langevin-gamma": 20.0, "temperature": 300.0,
       "ele-alpha": 0.0, "gap-bug": 0}


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _walker(n, seed, vcmd=False):
    ls = gen.walker_system(n, seed=seed, vcmd=vcmd)
    ls.settle_ids, ls.settle_d = ls.shake_trio_ids, ls.shake_trio_d
    ls.shake_trio_ids = np.zeros((0, 3), np.int32)
    ls.shake_trio_d = np.zeros((0, 3))
    return ls


def _solo(ls, cfg, steps):
    s = System(ls, cfg)
    s.run(steps)
    st = s.state()
    s.close()
    return st


def test_walkers_at_different_rebuild_phases_share_a_batch():
    """A walker at step 5 (rebuild interval 10) joins one at step 0: each
    rebuilds at its own phase and both equal their runs alone."""
    A, B = _walker(3000, 31), _walker(4500, 32)
    ca, cb = dict(CFG, **{"random-seed": 1}), dict(CFG, **{"random-seed": 2})
    ref_a = _solo(A, ca, 20)
    ref_b = _solo(B, cb, 25)
    say("phases", A.n_atoms, B.n_atoms)
    sa, sb = System(A, ca), System(B, cb)
    sb.run(5)
    b = Batch()
    b.add(sa)
    b.add(sb)
    b.run(20)
    for s, ref in ((sa, ref_a), (sb, ref_b)):
        st = s.state()
        assert all(np.array_equal(x, y) for x, y in zip(st, ref))
    b.close()
    sa.close()
    sb.close()


def test_split_tail_system_in_a_batch_equals_its_run_alone():
    """A system large enough for the pair kernel's split tail (the last 4,096
    of its 8,192+ cells as half-record waves) batched with a small walker (one
    wave per cell): both equal their runs alone bitwise over 12 steps (a grid
    rebuild included), in every block order of the batch."""
    big = gen.build_system(23000, seed=33)
    big.settle_ids, big.settle_d = big.shake_trio_ids, big.shake_trio_d
    big.shake_trio_ids = np.zeros((0, 3), np.int32)
    big.shake_trio_d = np.zeros((0, 3))
    small = _walker(3000, 34)
    cb, cs = dict(CFG, **{"random-seed": 3}), dict(CFG, **{"random-seed": 4})
    ref_big = _solo(big, cb, 12)
    ref_small = _solo(small, cs, 12)
    say("split tail in a batch", big.n_atoms, small.n_atoms)
    assert big.n_atoms // 8 >= 2 * 4096
    for order in ("system", "round-robin", "xcd", "xcd-parts"):
        sbig, ssmall = System(big, cb), System(small, cs)
        b = Batch()
        b.add(sbig)
        b.add(ssmall)
        b.set_pair_shape(4, order)
        b.run(12)
        for s, ref in ((sbig, ref_big), (ssmall, ref_small)):
            assert all(np.array_equal(x, y) for x, y in zip(s.state(), ref)), order
        b.close()
        sbig.close()
        ssmall.close()


def test_faulted_walker_stands_still_and_the_others_go_on():
    """One member's coordinates put two oxygens 0.05 A apart: its pair rows
    exceed the fixed-point range, it is flagged and stands still (step 0),
    while its batch partner runs the 20 steps exactly as alone."""
    A, B = _walker(3000, 33), _walker(3200, 34)
    ca, cb = dict(CFG, **{"random-seed": 3}), dict(CFG, **{"random-seed": 4})
    ref_a = _solo(A, ca, 20)
    sa, sb = System(A, ca), System(B, cb)
    crd = B.crd.copy()
    crd[3] = crd[0] + np.array([0.05, 0.0, 0.0])   # atom 3 is the next water's oxygen
    sb.set_state(crd, B.vel)
    b = Batch()
    b.add(sa)
    b.add(sb)
    with pytest.raises(EngineError) as e:
        b.run(20)
    say("fault contained:", e.value)
    assert e.value.code == OG_E_NAN
    assert sa.status() == 0 and sb.status() == OG_E_NAN
    assert sb.step() == 0 and sa.step() == 20
    assert all(np.array_equal(x, y) for x, y in zip(sa.state(), ref_a))
    b.close()
    sa.close()
    sb.close()


def test_set_state_mid_interval_rebuilds_the_grid():
    """og_set_state at step 5 with a rebuild interval of 10: the next step's
    forces are those of a system created at the new coordinates."""
    ls = _walker(3000, 35)
    cfg = dict(CFG, **{"integrator": "leapfrog-presto", "constraint": "none"})
    s = System(ls, cfg)
    s.run(5)
    rng = np.random.default_rng(7)
    crd = ls.crd + rng.normal(scale=0.05, size=ls.crd.shape)
    s.set_state(crd, ls.vel)
    s.run(1)
    w = s.work()
    s.close()
    import copy
    ls2 = copy.copy(ls)
    ls2.crd = crd
    f = System(ls2, cfg)
    w2, _ = f.energy_force(rebuild=True)
    f.close()
    say("set_state: max |dw| %.3e" % np.max(np.abs(w - w2)))
    assert np.array_equal(w, w2)


def test_serving_stream_with_resubmission_under_load():
    """A seeded stream in the shape of the serving workload: walkers of
    2,000-50,000 atoms (log-uniform), 280-360 K, arriving over successive
    pumps; segment lengths that are not multiples of the rebuild interval,
    and two walkers resubmitted as soon as their first segment ends, while
    others are still running.  Every walker ends bitwise where its run
    alone of the same total steps ends."""
    rng = np.random.default_rng(17)
    n_req = 6
    sizes = np.exp(rng.uniform(np.log(2000), np.log(50000), n_req)).astype(int)
    steps = [125, 240, 95, 180, 60, 150]
    temps = rng.uniform(280.0, 360.0, n_req)
    arrive = [0, 0, 1, 3, 4, 6]          # pump index at which each request arrives
    extra = {0: 35, 2: 55}               # resubmissions (walker index -> steps)
    ws = [_walker(int(n), 700 + i) for i, n in enumerate(sizes)]
    cfgs = [dict(CFG, **{"langevin-gamma": 1.0, "temperature": float(temps[i]), "random-seed": 40 + i})
            for i in range(n_req)]
    say("serving stream", [w.n_atoms for w in ws])
    srv = Server(16 << 30, chunk_steps=20)
    req, wid = {}, {}
    pumps, nxt = 0, 0
    resub = {}
    while True:
        while nxt < n_req and arrive[nxt] <= pumps:
            r, w = srv.submit_new(ws[nxt], cfgs[nxt], steps[nxt])
            req[r], wid[nxt] = nxt, w
            nxt += 1
        active, queued, _ = srv.pump()
        pumps += 1
        for i, k in extra.items():
            if i in wid and i not in resub:
                try:
                    srv.result([r for r, j in req.items() if j == i][0])
                except EngineError:
                    continue            # not finished yet
                resub[i] = srv.submit_next(wid[i], k)
        if nxt == n_req and active == 0 and queued == 0 and len(resub) == len(extra):
            break
        assert pumps < 1000
    finals = {}
    for i in range(n_req):
        rid = resub.get(i, [r for r, j in req.items() if j == i][0])
        finals[i] = srv.result(rid)[0]
    srv.close()
    for i in range(n_req):
        ref = _solo(ws[i], cfgs[i], steps[i] + extra.get(i, 0))
        assert np.array_equal(finals[i], ref[0]), i
    say("  %d walkers, %d pumps: every final state equals its solo run" % (n_req, pumps))


def test_64_walker_batch_in_500_step_graphs():
    """The walker workload's shape: 64 water + chain walkers of 3,000-30,000
    atoms (log-uniform), each with the 8-window VcMD bias, one batch run as
    a single 500-step graph (position Langevin at 20/ps from the lattice).
    No member raises a flag, and three sampled members equal their runs
    alone."""
    sizes = gen.walker_sizes(64)
    ws = [_walker(int(n), 5000 + i, vcmd=True) for i, n in enumerate(sizes)]
    cfgs = [dict(CFG, **{"random-seed": 100 + i, "extended-ensemble": "vcmd", "enhance-sigma": 0.5,
                         "enhance-recovery-coef": 1.0}) for i in range(64)]
    say("64 walkers, %d atoms" % sum(w.n_atoms for w in ws))
    ss = [System(w, c) for w, c in zip(ws, cfgs)]
    b = Batch()
    for s in ss:
        b.add(s)
    b.set_graph_steps(500)
    b.run(500)
    assert all(s.status() == 0 and s.step() == 500 for s in ss)
    picks = [0, 31, 63]
    got = {k: ss[k].state() for k in picks}
    b.close()
    for s in ss:
        s.close()
    for k in picks:
        ref = _solo(ws[k], cfgs[k], 500)
        assert all(np.array_equal(x, y) for x, y in zip(got[k], ref)), k
    say("  500 steps, no flags; walkers %s equal their runs alone" % picks)
