"""Continuous batching (og_server_*): a segment's result is the same alone
and in a mix of walkers admitted at different times, and the same as the
plain single-system run; a resubmitted walker continues from its device
state.  Fixed-point forces and per-walker Philox noise make these bitwise."""
import sys

import numpy as np
import pytest

from omegagene_amd import gen
from omegagene_amd.engine import Server, System

pytestmark = pytest.mark.gpu

CFG = {"cutoff": 12.0, "nsgrid-cutoff": 13.0, "nsgrid-update-intvl": 10, "time-step": 2.0, "integrator": "langevin",
       "constraint": "shake-settle", "const-tolerance": 1e-8, "langevin-gamma": 1.0, "temperature": 300.0,
       "ele-alpha": 0.0}


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _walker(n, seed):
    ls = gen.walker_system(n, seed=seed, vcmd=False)
    ls.settle_ids, ls.settle_d = ls.shake_trio_ids, ls.shake_trio_d
    ls.shake_trio_ids = np.zeros((0, 3), np.int32)
    ls.shake_trio_d = np.zeros((0, 3))
    return ls


def test_segment_same_alone_in_mix_and_single_path():
    ws = [_walker(n, 900 + i) for i, n in enumerate([3000, 4500, 6000])]
    cfgs = [dict(CFG, **{"random-seed": 11 + i, "temperature": 290.0 + 20 * i}) for i in range(3)]
    say("serving isolation", [w.n_atoms for w in ws])
    # the plain path: 55 + 20 steps of walker 0
    s = System(ws[0], cfgs[0])
    s.run(55)
    ref55 = s.state()[0]
    s.run(20)
    ref75 = s.state()[0]
    s.close()
    # alone in a server
    srv = Server(8 << 30, chunk_steps=10)
    r0, w0 = srv.submit_new(ws[0], cfgs[0], 55)
    while srv.pump()[0]:
        pass
    alone55 = srv.result(r0)[0]
    srv.close()
    # in a mix: different lengths, one walker arrives after two pumps, then a resubmission
    srv = Server(8 << 30, chunk_steps=10)
    ra, wa = srv.submit_new(ws[0], cfgs[0], 55)
    rb, wb = srv.submit_new(ws[1], cfgs[1], 40)
    srv.pump()
    srv.pump()
    rc, wc = srv.submit_new(ws[2], cfgs[2], 70)
    while srv.pump()[0]:
        pass
    crd_a, _, rows_a, t_a = srv.result(ra)
    assert len(rows_a) == 55
    rn = srv.submit_next(wa, 20)
    while srv.pump()[0]:
        pass
    crd_a2 = srv.result(rn)[0]
    st = srv.stats()
    say("  steps served %d, walkers %d, bytes %.1f MB" % (st["steps_done"], st["n_walkers"], st["bytes_used"] / 1e6))
    srv.close()
    assert np.array_equal(alone55, ref55)
    assert np.array_equal(crd_a, ref55)
    assert np.array_equal(crd_a2, ref75)
    assert st["steps_done"] == 55 + 40 + 70 + 20


def test_admission_footprint_bounds_the_allocation():
    """The footprint the server admits against (og_system_footprint, from the
    description alone) is an upper estimate of the device memory a created
    system holds, within a factor of two, from walker sizes to 1.2 M atoms."""
    from omegagene_amd.engine import footprint
    cases = [(_walker(3000, 11), CFG), (_walker(30000, 12), CFG), (_walker(99200, 13), CFG)]
    ls2, cfg2 = gen.config2()
    cases.append((ls2, dict(CFG, **{"cutoff": 10.0, "nsgrid-cutoff": 11.0, "ele-alpha": 0.1})))
    for ls, cfg in cases:
        est = footprint(ls, cfg)
        s = System(ls, cfg)
        used = s.device_bytes()
        s.close()
        say("footprint %d atoms: estimate %.1f MB, allocated %.1f MB" % (ls.n_atoms, est / 1e6, used / 1e6))
        assert used <= est <= 2 * used
