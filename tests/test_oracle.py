"""The oracle (oracle/*.c) pinned against everything the reference holds for
this path: the ala3 sample log (samples/ala3/log03_md.txt:99-135), MiniCell
grids and cell-pair lists from the reference sources, reference all-pairs
runs (one-step gradients, a 20-step SHAKE trajectory), and closed forms."""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import ala3_launchset, cls_from_fixture, ele_scale, load_npz, oracle_forces, rms_rel, wrap_inbox
from omegagene_amd import gen


def fmt(x):
    return "%.10e" % x


def test_ala3_step0_vdw_ele_match_log():
    ls, exp = ala3_launchset()
    zd = oracle.zd_setup(12.0, 0.0)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    tab = ls.nb15off_table()
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    work, en, npair = oracle.allpairs_fp32(zd, L, crd, ls.atom_type, lj6, lj12, q, tab)
    ex = ls.excess_pairs()
    assert len(ex) == 6435  # log03_md.txt:79 "excess set 6435"
    eex = oracle.excess(zd, L, ex, crd, q, np.zeros_like(work))
    eself, _ = oracle.self_energy(zd, q, ex)
    assert fmt(en[0]) == fmt(exp["vdw"])                     # VDW 4.1648848106e+03
    assert fmt(en[1] + eex + eself) == fmt(exp["ele"])      # Ele -2.9051201002e+04


def test_zd_constants_and_self_energy_kat():
    ls, exp = ala3_launchset()
    zd = oracle.zd_setup(12.0, 0.0)
    assert "%.15g" % zd.fcoeff == "%.15g" % exp["fcoeff"]
    assert "%.15g" % zd.bcoeff == "%.15g" % exp["bcoeff"]
    assert zd.zcore == exp["zmcore"] == 0.125
    eself, _ = oracle.self_energy(zd, ls.charge.astype(np.float32), ls.excess_pairs())
    assert "%.15g" % eself == "%.15g" % exp["self_sum"]


@pytest.mark.parametrize("name", ["w1500", "cfg0", "chain", "chain200", "rc11"])
def test_oracle_grid_matches_reference_minicell(name):
    d = load_npz("minicell_%s.npz" % name)
    n = len(d["crd"])
    o = oracle.Oracle(n, d["box"], float(d["rcl"]))
    c = o.build(d["crd"])
    assert np.array_equal(c.atomids, d["atomids"])
    assert np.array_equal(c.crd, d["grid_crd"])
    assert np.array_equal(c.idx_xy_head_cell, d["idx_xy_head_cell"])
    cps, head = o.enumerate(d["nb15off"], gap_bug=True)
    assert np.array_equal(head, d["idx_head_cell_pairs"])
    assert np.array_equal(cps.view(np.int32).reshape(-1, 5), d["cell_pairs"])


@pytest.mark.parametrize("tag,alpha", [("a0", 0.0), ("a1", 0.1)])
def test_oracle_gradient_matches_reference_one_step(tag, alpha):
    d = load_npz("wons_grad_%s.npz" % tag)
    ls = cls_from_fixture(d)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    zd = oracle.zd_setup(12.0, alpha)
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    work, en, _ = oracle.allpairs_fp32(zd, L, crd, ls.atom_type, lj6, lj12, q, ls.nb15off_table())
    ex = ls.excess_pairs()
    eb = oracle.bonds(L, ls.bond_ids, ls.bond_k, ls.bond_r0, crd, work)
    eex = oracle.excess(zd, L, ex, crd, q, work, bug_compat=True)
    eself, _ = oracle.self_energy(zd, q, ex)
    ref = d["work"]
    assert rms_rel(work, ref) < 1e-11
    assert np.max(np.abs(work - ref)) < 1e-8 * np.max(np.abs(ref))
    log0 = d["log"][0]  # step, pot, kin, bond, vdw, ele, T
    assert fmt(en[0]) == fmt(log0[4])
    assert fmt(en[1] + eex + eself) == fmt(log0[5])
    assert abs(eb - log0[3]) <= 1e-9 * max(1.0, abs(log0[3]))


def _oracle_leapfrog_shake(ls, n_steps, dt=2.0, tol=1e-10):
    """DynamicsModePresto::calc_in_each_step (DynamicsMode.cpp:468-560) with the
    oracle's all-pairs forces, SHAKE, no thermostat, no COM cancel."""
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    zd = oracle.zd_setup(12.0, 0.0)
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    tab = ls.nb15off_table()
    ex = ls.excess_pairs()
    minv = (1.0 / ls.mass.astype(np.float32).astype(np.float64)).astype(np.float32)
    m32 = ls.mass.astype(np.float32)
    d2 = ls.shake_trio_d ** 2
    eself, _ = oracle.self_energy(zd, q, ex)
    vel_next = ls.vel.copy()
    rows = []
    for step in range(n_steps):
        work, en, _ = oracle.allpairs_fp32(zd, L, crd, ls.atom_type, lj6, lj12, q, tab)
        eex = oracle.excess(zd, L, ex, crd, q, work)
        vel = vel_next
        vel_next = oracle.update_velocities(dt, vel, work, minv)
        crd_prev = crd
        crd = oracle.update_coordinates(dt, crd_prev, vel_next)
        crd, nf = oracle.shake_trio(L, ls.shake_trio_ids, d2, crd, crd_prev, minv, tol, 1000)
        assert nf == 0
        vel_next = oracle.velocity_from_crd(dt, crd, crd_prev)
        crd = oracle.revise_pbc(L, lb, crd)
        ke = oracle.kinetic(m32, vel, vel_next)
        rows.append([step, en[0] + en[1] + eex + eself, ke, en[0], en[1] + eex + eself])
    return crd, vel_next, np.array(rows)


def test_oracle_md_matches_reference_trajectory():
    d = load_npz("wons_traj.npz")
    ls = cls_from_fixture(d)
    crd, vel, rows = _oracle_leapfrog_shake(ls, 20)
    log = d["log"]  # step, pot, kin, bond, vdw, ele, T
    for r, lg in zip(rows, log[:20]):
        assert abs(r[1] - lg[1]) <= 2e-9 * abs(lg[5]) + 1e-8, (r, lg)   # potential
        assert abs(r[2] - lg[2]) <= 1e-9 * abs(lg[2]), (r, lg)          # kinetic
        assert fmt(r[3]) == fmt(lg[4])                                    # VDW at print precision
    assert np.max(np.abs(crd - d["crd"])) < 1e-9
    assert np.max(np.abs(vel - d["vel"])) < 1e-11


# ---------------------------------------------------------------- closed forms

def test_zd_energy_and_force_vanish_at_cutoff():
    zd = oracle.zd_setup(12.0, 0.0)
    for r in (12.0 * (1 - 1e-6), 11.9999):
        n, ev, ee, w = oracle.pair_fp32(zd, np.array([r, 0, 0]), 0.0, 0.0, 1.0, -1.0)
        assert n == 1
        assert abs(ee) < 1e-3 * 332.06378 / 12.0 * 1e-1
        assert abs(w[0]) < 1e-4 * 332.06378 / 144.0
    n, *_ = oracle.pair_fp32(zd, np.array([12.0, 0, 0]), 1.0, 1.0, 1.0, 1.0)
    assert n == 0  # r >= cutoff is skipped (ForceField.cpp:368)


def test_pair_two_particle_closed_form():
    zd = oracle.zd_setup(12.0, 0.0)
    r, a6, a12, q1, q2 = 3.3, 600.0, 6e5, -0.8, 0.4
    n, ev, ee, w = oracle.pair_fp32(zd, np.array([0.0, 0.0, r]), a6, a12, q1, q2)
    cc = q1 * q2 * 332.06378
    rc = 12.0
    assert ev == pytest.approx(a12 / r ** 12 - a6 / r ** 6, rel=2e-6)
    assert ee == pytest.approx(cc * (1 / r - 1 / rc - 0.5 / rc ** 3 * rc ** 2 + 0.5 / rc ** 3 * r ** 2), rel=2e-6)
    # dE/dr * (x/r): derivative of the same closed form
    dEdr = -12 * a12 / r ** 13 + 6 * a6 / r ** 7 + cc * (-1 / r ** 2 + r / rc ** 3)
    assert w[2] == pytest.approx(dEdr, rel=2e-6)
    assert w[0] == 0 and w[1] == 0


def test_fp64_pair_gradient_finite_difference():
    ls = gen.build_system(120, seed=3, L=16.0)  # tiny dense box, rc 7 / list 8
    ls.crd[:] = np.mod(ls.crd, 16.0)
    zd = oracle.zd_setup(7.0, 0.1)
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)

    def energy_and_grad(crd):
        o = oracle.Oracle(ls.n_atoms, ls.box, 8.0)
        o.build(crd)
        cps, _ = o.enumerate(ls.nb15off_table(), gap_bug=False)
        w, en, _ = o.pairs_fp64(zd, cps, crd, ls.atom_type, lj6, lj12, q)
        return en[0] + en[1], w

    e0, w0 = energy_and_grad(ls.crd)
    h = 1e-5
    for a in (0, 7, 200):
        for d_ in range(3):
            cp, cm = ls.crd.copy(), ls.crd.copy()
            cp[a, d_] += h
            cm[a, d_] -= h
            fd = (energy_and_grad(cp)[0] - energy_and_grad(cm)[0]) / (2 * h)
            assert fd == pytest.approx(w0[a, d_], rel=1e-5, abs=1e-6)


def test_cell_list_covers_every_pair_once():
    d = load_npz("minicell_cfg0.npz")
    crd, box, rc = d["crd"], d["box"], 12.0
    L = box[0]
    o = oracle.Oracle(len(crd), box, 13.0)
    c = o.build(crd)
    cps, _ = o.enumerate(d["nb15off"], gap_bug=True)
    seen = {}
    ids = c.atomids
    gx = c.crd.astype(np.float32)
    for rec in cps:
        c1, c2, img, m0, m1 = int(rec[0]), int(rec[1]), int(rec[2]), int(rec[3]), int(rec[4])
        mask = (m0 & 0xffffffff) | ((m1 & 0xffffffff) << 32)
        shift = np.array([(-L if img & 1 else L if img & 2 else 0.0), (-L if img & 4 else L if img & 8 else 0.0),
                          (-L if img & 16 else L if img & 32 else 0.0)])
        a1 = ids[c1 * 8:c1 * 8 + 8]
        a2 = ids[c2 * 8:c2 * 8 + 8]
        for j in range(8):
            for i in range(8):
                if (mask >> (j * 8 + i)) & 1:
                    continue
                r = np.linalg.norm(gx[c1 * 8 + i].astype(np.float64) - (gx[c2 * 8 + j] + shift))
                if r >= rc:
                    continue
                key = (min(a1[i], a2[j]), max(a1[i], a2[j]))
                seen[key] = seen.get(key, 0) + 1
    assert max(seen.values()) == 1
    # brute force over all pairs (float coordinates, minimum image)
    x = gx[ids >= 0]
    order = ids[ids >= 0]
    pos = np.empty_like(x)
    pos[order] = x
    excl = set()
    for a, row in enumerate(d["nb15off"]):
        for b in row:
            if b < 0:
                break
            excl.add((min(a, b), max(a, b)))
    n_in = 0
    for a in range(len(pos)):
        dd = pos[a + 1:] - pos[a]
        dd -= L * np.round(dd / L)
        r = np.sqrt((dd.astype(np.float64) ** 2).sum(1))
        for b in np.nonzero(r < rc)[0] + a + 1:
            if (a, b) in excl:
                continue
            n_in += 1
            assert (a, b) in seen
    assert n_in > 0


def test_settle_matches_converged_shake():
    ls = gen.build_system(300, seed=9, L=21.0)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    rng = np.random.default_rng(0)
    prev = ls.crd.copy()
    cur = prev + rng.normal(scale=0.02, size=prev.shape)  # an unconstrained step
    m32 = ls.mass.astype(np.float32)
    minv = (1.0 / m32.astype(np.float64)).astype(np.float32)
    d2 = ls.shake_trio_d ** 2
    sh, nf = oracle.shake_trio(L, ls.shake_trio_ids, d2, cur, prev, minv, 1e-14, 1000)
    assert nf == 0
    st = oracle.settle(L, ls.shake_trio_ids, gen.D_OH, gen.D_HH, cur, prev, m32)
    assert np.max(np.abs(st - sh)) < 1e-9
    w = ls.shake_trio_ids
    for a, b, dist in ((0, 1, gen.D_OH), (1, 2, gen.D_HH), (0, 2, gen.D_OH)):
        r = np.linalg.norm(st[w[:, a]] - st[w[:, b]], axis=1)
        assert np.max(np.abs(r - dist)) < 1e-12
    # linear momentum of each water is untouched by the constraint forces
    mom0 = (m32[w][..., None] * (cur[w] - prev[w])).sum(1)
    mom1 = (m32[w][..., None] * (st[w] - prev[w])).sum(1)
    assert np.max(np.abs(mom0 - mom1)) < 1e-12


def test_langevin_zero_noise_zero_friction_is_verlet():
    rng = np.random.default_rng(1)
    n = 50
    x0 = rng.uniform(0, 10, (n, 3))
    x1 = x0 + rng.normal(scale=0.01, size=(n, 3))
    work = rng.normal(size=(n, 3))
    m = rng.uniform(1, 16, n).astype(np.float32)
    minv = (1.0 / m.astype(np.float64)).astype(np.float32)
    L = np.array([100.0, 100.0, 100.0], np.float32)
    dt = 2.0
    x2 = oracle.langevin_positions(dt / 2, 0.0, 300.0, L, x1, x0, work, m, minv, np.zeros((n, 3)))
    verlet = 2 * x1 - x0 + dt * dt * minv[:, None].astype(np.float64) * (-4.184e-4 * work)
    assert np.max(np.abs(x2 - verlet)) < 1e-12


# ---------------------------------------------------------------- VcMD bias

def test_vcmd_bias_closed_form_two_atoms():
    """Two single-atom groups: lambda is their distance; outside
    [min - sigma, max + sigma) the gradient of atom a is
    +k_B T * coef * excess * u and of atom b its negative, and the energy is
    coef * excess^2 (Extend.cpp:1035-1102; the energy/force pair of the
    reference is kept as is)."""
    L = np.array([40.0, 40.0, 40.0], np.float32)
    crd = np.array([[1.0, 2.0, 3.0], [1.0 + 6.0, 2.0 + 8.0, 3.0]])  # distance 10
    mass = np.array([12.0, 12.0])
    lo, hi = np.array([2.0, 4.0]), np.array([5.0, 7.0])
    T, sigma, coef = 300.0, 0.5, 2.0
    kT = (8.31451 / 4.184) * 1e-3 * T
    for vs, edge in ((0, 5.0 + 0.5), (1, 7.0 + 0.5)):
        work = np.zeros((2, 3))
        crdg = crd.copy()
        e, lam, inr, cand = oracle.vcmd_bias(L, [0], [1], crdg, crd, mass, lo, hi, vs, sigma, coef, T, work)
        # lambda carries the float group inverse mass (1/12 rounded to float)
        assert lam == pytest.approx(10.0, rel=1e-7)
        ex = lam - edge
        u = (crd[0] - crd[1]) / 10.0
        assert e == pytest.approx(coef * ex * ex, rel=1e-12)
        assert np.allclose(work[0], kT * coef * ex * u, rtol=1e-6)  # float m/M_group
        assert np.allclose(work[1], -work[0], rtol=1e-12)
        assert not inr
        # from state 0 the neighbour 1 is a candidate once lambda >= its min (4); state 1 has no upper neighbour
        assert cand == ((0, 1) if vs == 0 else (1, -1))
    # inside the widened range: no bias, and counted as a visit only inside [min, max)
    work = np.zeros((2, 3))
    crd2 = np.array([[0.0, 0.0, 0.0], [5.2, 0.0, 0.0]])
    e, lam, inr, cand = oracle.vcmd_bias(L, [0], [1], crd2.copy(), crd2, mass, lo, hi, 0, sigma, coef, T, work)
    assert e == 0.0 and not work.any() and not inr and cand == (0, 1)   # 5.2 >= min of state 1
    e, lam, inr, cand = oracle.vcmd_bias(L, [0], [1], crd2.copy(), crd2, mass, lo, hi, 1, sigma, coef, T, work)
    assert inr and cand == (1, -1)                                        # 5.2 >= max of state 0


def test_vcmd_group_unwrap_and_mass_centre():
    """A group straddling the periodic boundary: the mass centre follows the
    unwrapped coordinates kept between calls, so lambda is continuous."""
    Lf = 20.0
    L = np.array([Lf, Lf, Lf], np.float32)
    mass = np.array([1.0, 3.0, 2.0])
    crd = np.array([[9.5, 0.0, 0.0], [-9.5, 0.0, 0.0], [0.0, 0.0, 0.0]])   # atoms 0,1 are 1 A apart across x
    crdg = np.array([[9.5, 0.0, 0.0], [10.5, 0.0, 0.0], [0.0, 0.0, 0.0]])  # previous unwrapped positions
    work = np.zeros((3, 3))
    lo, hi = np.array([0.0]), np.array([100.0])
    e, lam, inr, cand = oracle.vcmd_bias(L, [0, 1], [2], crdg, crd, mass, lo, hi, 0, 0.0, 1.0, 300.0, work)
    com_a = (9.5 * 1.0 + 10.5 * 3.0) / 4.0           # 10.25 -> minimum image -9.75 from atom 2
    assert crdg[1, 0] == 10.5
    assert lam == pytest.approx(abs(com_a - Lf), rel=1e-6)
    assert inr and e == 0.0


def test_fp64_cell_list_equals_brute_force():
    """The FP64 error reference (orc_pairs_celllist_fp64) against a numpy
    all-pairs minimum-image sum on a small box where one cell pair is listed
    under several z images (so an image taken from the coordinates instead
    of the cell pair would double-count)."""
    ls = gen.build_system(400, seed=12, L=24.5)
    zd = oracle.zd_setup(12.0, 0.0)
    o = oracle.Oracle(ls.n_atoms, ls.box, 13.0)
    c = o.build(ls.crd)
    cps, _ = o.enumerate(ls.nb15off_table(), gap_bug=False)
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    w64, e64, n64 = o.pairs_fp64(zd, cps, ls.crd, ls.atom_type, lj6, lj12, q)
    # brute force, FP64, every unexcluded pair within the cutoff once
    L = float(np.float32(ls.box[0]))
    x = ls.crd
    n = ls.n_atoms
    tab = ls.nb15off_table()
    excl = np.zeros((n, n), bool)
    for i in range(n):
        excl[i, tab[i][tab[i] >= 0]] = True
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d * d).sum(-1)
    iu = np.triu_indices(n, 1)
    ok = (r2[iu] < 144.0) & ~excl[iu]
    i, j = iu[0][ok], iu[1][ok]
    r = np.sqrt(r2[i, j])
    t = ls.atom_type[i] * ls.n_lj_types + ls.atom_type[j]
    a6, a12 = lj6.reshape(-1)[t].astype(np.float64), lj12.reshape(-1)[t].astype(np.float64)
    cc = q[i].astype(np.float64) * q[j].astype(np.float64) * 332.06378
    ri = 1.0 / r
    g = ri ** 2 * (-12 * a12 * ri ** 12 + 6 * a6 * ri ** 6) - cc * (ri ** 3 - 1.0 / 12.0 ** 3)
    wb = np.zeros((n, 3))
    np.add.at(wb, i, g[:, None] * d[i, j])
    np.add.at(wb, j, -g[:, None] * d[i, j])
    assert n64 == len(r)
    assert rms_rel(w64, wb) < 1e-12
    assert e64[0] == pytest.approx(np.sum(a12 * ri ** 12 - a6 * ri ** 6), rel=1e-12)
    # virial: sum over pairs of g r^2 (closed form per pair, brute force)
    assert e64[2] == pytest.approx(np.sum(g * r * r), rel=1e-12)


def test_settle_reports_waters_without_rigid_solution():
    """A water pushed so far that no rotation of the rigid model reaches its
    displaced positions is left unconstrained and counted, never NaN."""
    L = np.array([30.0, 30.0, 30.0], np.float32)
    dOH, dHH = gen.D_OH, gen.D_HH
    prev = np.array([[0.0, 0.0, 0.0], [dOH, 0.0, 0.0], [dOH * np.cos(np.radians(104.52)), dOH * np.sin(np.radians(104.52)), 0.0]])
    cur = prev.copy()
    cur[1, 2] += 3.0    # H1 thrown 3 A out of the plane
    ids = np.array([[0, 1, 2]], np.int32)
    mass = np.array([15.9994, 1.008, 1.008], np.float32)
    c = np.ascontiguousarray(cur.copy())
    f = oracle.lib().orc_settle
    f.restype = ctypes.c_int
    n_fail = f(oracle._p(L, ctypes.c_float), ctypes.c_int(1), oracle._p(ids, ctypes.c_int), ctypes.c_double(dOH),
               ctypes.c_double(dHH), oracle._p(c, ctypes.c_double), oracle._p(np.ascontiguousarray(prev), ctypes.c_double),
               oracle._p(mass, ctypes.c_float))
    assert n_fail == 1
    assert np.array_equal(c, cur)
    # a small displacement is solved exactly
    cur2 = prev + 0.01
    cur2[2, 0] += 0.02
    c2 = oracle.settle(L, ids, dOH, dHH, cur2, prev, mass)
    assert abs(np.linalg.norm(c2[1] - c2[0]) - dOH) < 1e-12 and abs(np.linalg.norm(c2[2] - c2[1]) - dHH) < 1e-12


def test_settle_keeps_each_atom_in_its_periodic_frame():
    """A water whose hydrogen is wrapped to the other side of the box: SETTLE
    moves every atom by a small correction in its own frame (no jump of one
    box length), and the constrained geometry holds under the minimum image."""
    Lf = 20.0
    L = np.array([Lf, Lf, Lf], np.float32)
    dOH, dHH = gen.D_OH, gen.D_HH
    th = np.radians(104.52)
    O = np.array([9.7, 1.0, 1.0])
    prev = np.array([O, O + [dOH, 0, 0], O + [dOH * np.cos(th), dOH * np.sin(th), 0]])
    prev[1, 0] -= Lf                       # H1 stored wrapped: x = -9.54 instead of 10.46
    rng = np.random.default_rng(3)
    cur = prev + rng.normal(scale=0.02, size=(3, 3))
    mass = np.array([15.9994, 1.008, 1.008], np.float32)
    new = oracle.settle(L, np.array([[0, 1, 2]], np.int32), dOH, dHH, cur, prev, mass)
    assert np.max(np.abs(new - cur)) < 0.1
    d = lambda a, b: np.linalg.norm((new[a] - new[b]) - Lf * np.round((new[a] - new[b]) / Lf))
    assert abs(d(0, 1) - dOH) < 1e-12 and abs(d(0, 2) - dOH) < 1e-12 and abs(d(1, 2) - dHH) < 1e-12


def ref_two_pass_image(v, Lf):
    """PBC::diff_crd_minim_image (PBC.cpp:62-82) as written: two truncating
    passes with float L/2 and its float reciprocal."""
    L32 = np.float32(Lf)
    Lhi = np.float32(np.float32(1.0) / np.float32(L32 * np.float32(0.5)))
    for _ in range(2):
        v = v - float(L32 * np.float32(int(v * float(Lhi))))
    return v


def straddling_water(Lf, gap, dOH, dHH):
    """A rigid water in the xz plane (so O, H1, H2 share y up to gap) with O
    just below the top y face and H1 wrapped just above the bottom one: the
    H1 - O difference in y is gap - L."""
    th = np.radians(104.52)
    O = np.array([0.4 * Lf, Lf - gap / 2, 0.6 * Lf])
    w = np.array([O, O + [dOH, gap, 0.0], O + [dOH * np.cos(th), 0.0, dOH * np.sin(th)]])
    w[1, 1] -= Lf
    assert 0.0 < w[1, 1] < Lf and 0.0 < w[0, 1] < Lf
    return w


def test_minimum_image_at_one_box_length_and_settle_across_the_face():
    """Two atoms a hair apart in y across the box face (difference within
    float rounding of -L): the reference's two-pass minimum image swaps -L and
    +L and returns a vector one box long; the oracle (like the device) takes
    one nearest-image step more there, so SETTLE with such a water as its
    reference geometry keeps it rigid and gives the positions of the same water
    moved off the face, and the excluded-pair and bond terms see the short
    vector.  Any other difference keeps the reference's result."""
    Lf, gap = 30.0, 2e-8
    dOH, dHH = gen.D_OH, gen.D_HH
    prev = straddling_water(Lf, gap, dOH, dHH)
    v = prev[1, 1] - prev[0, 1]
    assert abs(ref_two_pass_image(v, Lf)) > 0.75 * Lf   # the input does hit the reference's cycle
    for u in (v + 1e-3, 7.3, -14.9, 15.1, -44.0):     # elsewhere the reference's passes already give |d| <= L/2
        assert abs(ref_two_pass_image(u, Lf)) <= 0.5 * Lf + 1e-6
    L = np.array([Lf] * 3, np.float32)
    mass = np.array([15.9994, 1.008, 1.008], np.float32)
    ids = np.array([[0, 1, 2]], np.int32)
    rng = np.random.default_rng(5)
    step = rng.normal(scale=0.01, size=(3, 3))
    new = oracle.settle(L, ids, dOH, dHH, prev + step, prev, mass)
    d = lambda x, a, b: np.linalg.norm((x[a] - x[b]) - Lf * np.round((x[a] - x[b]) / Lf))
    assert abs(d(new, 0, 1) - dOH) < 1e-12 and abs(d(new, 0, 2) - dOH) < 1e-12 and abs(d(new, 1, 2) - dHH) < 1e-12
    # the same water half a box lower (no atom near a face): the same correction
    sh = np.array([0.0, Lf / 2, 0.0])
    moved = oracle.settle(L, ids, dOH, dHH, prev - sh + step, prev - sh, mass)
    diff = (new - (moved + sh)) - Lf * np.round((new - (moved + sh)) / Lf)
    assert np.max(np.abs(diff)) < 1e-9
    # bond and excluded-pair terms across the face: the short vector
    w = np.zeros((3, 3))
    eb = oracle.bonds(L, np.array([[0, 1]], np.int32), np.array([300.0]), np.array([dOH]), prev, w)
    assert eb < 1e-9 and np.max(np.abs(w)) < 1e-6


# ------------------------------------------------ reference runs of the step loop
# Each fixture is a run of the reference all-pairs build (oracle/_ref/omegagene_wons,
# compiled from the reference sources by oracle/Makefile; tests/golden/make_wons_fixtures.py):
# the per-step log and the final restart.  helpers.oracle_md restates the reference
# step loop with the oracle's functions; the restart must agree exactly.

from helpers import oracle_md  # noqa: E402


def _check_run(d, crd, vel, rows, n_rows):
    assert np.max(np.abs(crd - d["crd"])) <= 1e-12, np.max(np.abs(crd - d["crd"]))
    assert np.max(np.abs(vel - d["vel"])) <= 1e-12, np.max(np.abs(vel - d["vel"]))
    log = d["log"]
    for r, lg in zip(rows[:n_rows], log[:n_rows]):
        assert fmt(r[4]) == fmt(lg[4])                                      # VDW at print precision
        assert abs(r[1] - lg[1]) <= 1e-9 * abs(lg[5]) + 1e-7               # potential (printed to 1e-7)
        if np.isfinite(r[2]):
            assert abs(r[2] - lg[2]) <= 1e-9 * abs(lg[2]) + 1e-7           # kinetic


def test_oracle_langevin_matches_reference_run():
    """Position Langevin (SubBox.cpp:2077-2114, DynamicsMode.cpp:798-899) with the
    reference's noise stream (std::mt19937 + a fresh std::normal_distribution per
    draw, Random.cpp:10-43): pins orc_langevin_positions (friction and noise
    terms at gamma 10/ps), orc_langevin_velocity and orc_ref_normals."""
    d = load_npz("wons_langevin.npz")
    ls = cls_from_fixture(d)
    crd, vel, rows = oracle_md(ls, 3, integrator="langevin", extra_last=True,
                               langevin=dict(gamma_per_fs=10.0 / 1000.0, temperature=300.0, seed=1234))
    _check_run(d, crd, vel, rows, 4)
    # the noise is live: the same steps without it end elsewhere
    L, lb, ub = oracle.pbc_from_box(ls.box)
    assert oracle.ref_normals(1234, 0, 6).std() > 0.1


def test_oracle_shake_pair_matches_reference_run():
    """SHAKE pairs on a chain (ConstraintShake.cpp:18-49) with SHAKE trios on the
    waters, 10 leapfrog steps: pins orc_shake_pair."""
    d = load_npz("wons_shake_pair.npz")
    ls = cls_from_fixture(d)
    assert len(ls.shake_pair_ids) > 0
    crd, vel, rows = oracle_md(ls, 10, constraint="shake")
    _check_run(d, crd, vel, rows, 10)
    L = float(np.float32(ls.box[0]))
    p = ls.shake_pair_ids
    v = crd[p[:, 0]] - crd[p[:, 1]]
    v -= L * np.round(v / L)
    assert np.max(np.abs(np.linalg.norm(v, axis=1) - ls.shake_pair_d)) < 1e-6


def test_oracle_com_cancel_matches_reference_run():
    """--com-motion cancel on two atom groups (COMMotion.cpp:30-55): pins
    orc_cancel_com; every group's momentum is zero after the removal."""
    d = load_npz("wons_com.npz")
    ls = cls_from_fixture(d)
    g = dict(ls.groups)
    groups = [g["solute"], g["solvent"]]
    crd, vel, rows = oracle_md(ls, 10, constraint="shake", com_groups=groups)
    _check_run(d, crd, vel, rows, 10)
    rng = np.random.default_rng(2)
    v = rng.normal(size=(ls.n_atoms, 3))
    m32 = ls.mass.astype(np.float32)
    v2 = oracle.cancel_com(groups, m32, v)
    for grp in groups:
        mom = (m32[grp, None].astype(np.float64) * v2[grp]).sum(0)
        scale = (m32[grp, None].astype(np.float64) * np.abs(v[grp])).sum(0)
        assert np.all(np.abs(mom) <= 1e-6 * scale)   # float group mass (MmSystem.cpp:572-582)


def test_oracle_thermostat_matches_reference_run():
    """--thermostat scaling without constraints (ThermostatScaling::apply_thermostat,
    Thermostat.cpp:57-84), --berendsen-tau 0.1 ps clamped to the 2 fs step
    (SubBox.cpp:1674-1676) and heating 250 -> 300 K over 5 steps
    (SubBox.cpp:1745-1757): pins orc_thermostat_scaling.  With tau clamped to
    dt the scale is sqrt(T0/T): the logged temperature is the target."""
    d = load_npz("wons_thermo.npz")
    ls = cls_from_fixture(d)
    th = dict(tau_fs=100.0, temperature=300.0, temperature_init=250.0, heating_steps=5, d_free=ls.d_free(False))
    crd, vel, rows = oracle_md(ls, 10, thermostat=th)
    _check_run(d, crd, vel, rows, 10)
    T = d["log"][:10, 6]
    assert np.allclose(T, [250, 260, 270, 280, 290, 300, 300, 300, 300, 300], rtol=1e-9)


def test_thermostat_scaling_closed_form():
    """One scaling step: the on-step kinetic temperature of v + dt/2 a is
    rescaled by sqrt(1 + dt/tau (T0/T - 1)) (Thermostat.cpp:57-84)."""
    rng = np.random.default_rng(4)
    n = 64
    m = rng.uniform(1, 16, n).astype(np.float32)
    minv = (1.0 / m.astype(np.float64)).astype(np.float32)
    v = rng.normal(scale=0.01, size=(n, 3))
    w = rng.normal(size=(n, 3))
    dt, tau, T0 = 2.0, 8.0, 310.0
    coeff = oracle.temperature_coeff(3 * n - 3)
    vn, s = oracle.thermostat_scaling(dt, 1.0 / tau, T0, coeff, w, v, m, minv)
    a = -4.184e-4 * w * minv[:, None].astype(np.float64)
    T = coeff * np.sum(m[:, None].astype(np.float64) * (v + 0.5 * dt * a) ** 2)
    assert s == pytest.approx(np.sqrt(1 + dt / tau * (T0 / T - 1)), rel=1e-13)
    assert np.allclose(vn, (2 * s - 1) * v + s * dt * a, rtol=1e-12, atol=1e-18)
    assert oracle.thermostat_tau_inv(2.0, 100.0) == 0.5 and oracle.thermostat_tau_inv(2.0, 1.0) == 1.0


@pytest.mark.parametrize("fixture", ["wons_grad_a0", "ala3"])
def test_celllist_pair_loop_equals_allpairs(fixture):
    """The FP32 cell-list loop (orc_pairs_celllist, SubBox.cpp:956-1125 over the
    MiniCell list) against the all-pairs loop pinned by the reference runs
    (orc_pairs_allpairs_fp32): the same pair count exactly.  The cell list
    takes image shifts on float grid coordinates where the all-pairs loop takes
    a float minimum image, so a pair's FP32 terms can differ in their last
    bits (the vdw totals differ at 4e-8 relative on the first system): energies
    and gradient agree to the FP32 rounding scale."""
    if fixture == "ala3":
        ls, _ = ala3_launchset()
    else:
        ls = cls_from_fixture(load_npz(fixture + ".npz"))
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    zd = oracle.zd_setup(12.0, 0.0)
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    tab = ls.nb15off_table()
    wa, ea, na = oracle.allpairs_fp32(zd, L, crd, ls.atom_type, lj6, lj12, q, tab)
    o = oracle.Oracle(ls.n_atoms, ls.box, 13.0)
    c = o.build(crd)
    cps, _ = o.enumerate(tab, gap_bug=False)
    wg, ec, nc = o.pairs_fp32(zd, cps, ls.atom_type, lj6, lj12, q)
    wc = np.zeros_like(wa)
    m = c.atomids >= 0
    wc[c.atomids[m]] = wg[m]
    assert nc == na
    # each pair's FP32 terms agree to a few float ulps: the sums to 2e-7 of the
    # vdw total and 1e-8 of the sum of the ele term magnitudes
    assert ec[0] == pytest.approx(ea[0], rel=2e-7)
    assert abs(ec[1] - ea[1]) <= 1e-8 * ele_scale(ls, crd=crd)
    err = rms_rel(wc, wa)
    assert err < 1e-6, err   # measured 5.0e-7 on the 24.5 A box, where most pairs are image pairs


def test_excess_fixed_mode_closed_form():
    """orc_excess with the reference's alpha > 0 energy term fixed (bug_compat 0:
    -erf(ar)/r instead of ZeroMultipoleSum.cpp:214's -erf(ar)*r) against the
    closed form per excluded pair, E = cc (-erf(a r)/r + bcoeff r^2) with the
    float constants, and the bug-compatible mode against its own form."""
    import math
    zd = oracle.zd_setup(12.0, 0.1)
    L = np.array([40.0, 40.0, 40.0], np.float32)
    rng = np.random.default_rng(6)
    # excluded pairs at bonded distances (0.9-3 A), one of them across the box edge
    crd = np.zeros((40, 3))
    crd[0::2] = rng.uniform(0, 40, (20, 3))
    u = rng.normal(size=(20, 3))
    crd[1::2] = crd[0::2] + u / np.linalg.norm(u, axis=1)[:, None] * rng.uniform(0.9, 3.0, 20)[:, None]
    crd[1] = crd[0] + [39.0, 0.5, 0.0]          # minimum image: 1.1 A apart
    q = rng.uniform(-0.8, 0.8, 40).astype(np.float32)
    pairs = np.array([[i, i + 1] for i in range(0, 40, 2)], np.int32)
    for bug in (0, 1):
        w = np.zeros((40, 3))
        e = oracle.excess(zd, L, pairs, crd, q, w, bug_compat=bug)
        ref, mag = 0.0, 0.0
        for a, b in pairs:
            d = crd[a].astype(np.float32) - crd[b].astype(np.float32)
            d = d - L * np.round(d / L)
            r = float(np.linalg.norm(d.astype(np.float64)))
            cc = float(q[a]) * float(q[b]) * 332.06378
            er = math.erf(0.1 * r)
            t = cc * ((-er * r if bug else -er / r) + float(zd.bcoeff) * r * r)
            ref += t
            mag += abs(t)
        assert abs(e - ref) <= 2e-6 * mag, (bug, e, ref, mag)
    # the two modes differ by cc erf(ar) (r - 1/r) summed over the pairs
    w = np.zeros((40, 3))
    diff = oracle.excess(zd, L, pairs, crd, q, w, bug_compat=1) - oracle.excess(zd, L, pairs, crd, q, w, bug_compat=0)
    assert abs(diff) > 1e-3


def test_ala3_degrees_of_freedom_and_initial_kinetic_energy():
    """samples/ala3/log03_md.txt:123-125: d_free 12417 = 3N - 3 - (pairs + 3 trios
    + 6 quads) + 3 with a position restraint (DynamicsMode.cpp:61-65,
    Read.cpp:90-94), and the initial kinetic energy 3701.28839 of the .cls
    velocities (DynamicsMode.cpp:358-372)."""
    d = load_npz("ala3_step0.npz")
    _, exp = ala3_launchset()
    n = len(d["crd"])
    npair, ntrio, nquad = [int(x) for x in d["n_shake"]]
    dfree = 3 * n - 3 - (npair + 3 * ntrio + 6 * nquad) + (3 if bool(d["has_pos_rest"]) else 0)
    assert dfree == int(exp["d_free"]) == 12417
    m32 = d["mass"].astype(np.float32)
    ke = oracle.kinetic(m32, d["vel"], d["vel"])
    assert "%.8e" % ke == "%.8e" % exp["ke_init"]
    T = ke * 2 * 4184.0 / (8.31451 * dfree)
    assert T == pytest.approx(exp["t_init"], rel=1e-9)


# ------------------------------------------------- HPS + Debye-Hueckel pair form

def test_oracle_hpsdh_gradient_matches_reference_one_step():
    """One leapfrog step of coarse-grained chains with --nonbond
    hydrophobicity-scale-lj and --electrostatic debye-huckel in the reference
    all-pairs build: the oracle's gradient to 1e-11 RMS and its VDW / Ele to
    the printed digits."""
    from helpers import hps_dh_form, oracle_forces_allpairs
    d = load_npz("wons_hpsdh_grad.npz")
    ls = cls_from_fixture(d)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    work, evdw, eele, eb = oracle_forces_allpairs(ls, crd, oracle.zd_setup(20.0, 0.0), L, ff=hps_dh_form(ls, d))
    ref = d["work"]
    assert rms_rel(work, ref) < 1e-11
    assert np.max(np.abs(work - ref)) < 1e-8 * np.max(np.abs(ref))
    log0 = d["log"][0]
    assert fmt(evdw) == fmt(log0[4])
    assert fmt(eele) == fmt(log0[5])
    assert abs(eb - log0[3]) <= 1e-9 * max(1.0, abs(log0[3]))


def test_oracle_hpsdh_md_matches_reference_trajectory():
    """20 leapfrog steps (dt 5 fs) of the same form at the cg_q8 sample's
    ionic strength: per-step energies and the final restart."""
    from helpers import hps_dh_form, oracle_md
    d = load_npz("wons_hpsdh_traj.npz")
    ls = cls_from_fixture(d)
    crd, vel, rows = oracle_md(ls, 20, dt=5.0, cutoff=20.0, ff=hps_dh_form(ls, d))
    log = d["log"]
    for r, lg in zip(rows, log[:20]):
        assert abs(r[1] - lg[1]) <= 1e-9 * (abs(lg[4]) + abs(lg[5]) + abs(lg[3])) + 1e-8, (r, lg)
        assert abs(r[2] - lg[2]) <= 1e-9 * abs(lg[2]), (r, lg)
        assert fmt(r[4]) == fmt(lg[4])
    assert np.max(np.abs(crd - d["crd"])) < 1e-9
    assert np.max(np.abs(vel - d["vel"])) < 1e-11


def test_hpsdh_two_particle_closed_form():
    """Two beads: E_DH = 332.06378 q1 q2 exp(-kappa r) / (eps_r r) with
    kappa = sqrt(2 N_A e^2 I / (eps_0 eps_r k_B T)) 1e-10 per A; HPS: beyond
    the switch distance lambda times LJ, inside it LJ + (1 - lambda) eps.
    The Debye-Hueckel gradient is the reference's -E (1/r + kappa) d
    (ZeroMultipoleSum.cpp:453, r times the derivative); dh_grad_fix gives
    -E (1/r + kappa) d / r, the central difference of E."""
    eps_r, I, T = 85.0, 150.0, 300.0
    kappa = np.sqrt(2 * 6.0221413e23 * 1.60217657e-19 ** 2 * I / (8.85418782e-12 * eps_r * 1.380658e-23 * T)) * 1e-10
    hc = np.array([[6.0]], np.float32)
    hl = np.array([[0.3]], np.float32)
    a6, a12 = 4 * 0.2 * 5.5 ** 6, 4 * 0.2 * 5.5 ** 12
    zd = oracle.zd_setup(20.0, 0.0)
    for fix in (False, True):
        ff = oracle.ff_setup(hc, hl, 0.2, eps_r, I, T, dh_grad_fix=fix)
        assert ff.debye_inv == pytest.approx(kappa, rel=1e-6)
        for r in (5.0, 6.5, 12.0):
            d12 = np.array([r, 0.0, 0.0], np.float32)
            n, ev, ee, w = oracle.pair_ff_fp32(zd, ff, d12, 0, a6, a12, 1.0, -1.0)
            lj = a12 / r ** 12 - a6 / r ** 6
            e_ref = -332.06378 * np.exp(-kappa * r) / (eps_r * r)
            assert ee == pytest.approx(e_ref, rel=2e-6)
            assert ev == pytest.approx(0.3 * lj if r > 6.0 else lj + 0.7 * 0.2, rel=2e-6, abs=1e-6)
            # the electrostatic part of the x-gradient
            glj = (-12 * a12 / r ** 14 + 6 * a6 / r ** 8) * (0.3 if r > 6.0 else 1.0)
            g_ele = (float(w[0]) - np.float32(glj) * r) / r
            dE = -e_ref * (1 / r + kappa)   # dE/dr
            assert g_ele == pytest.approx(dE / r if fix else dE, rel=1e-4)
        n, _, _, _ = oracle.pair_ff_fp32(zd, ff, np.array([20.0, 0, 0], np.float32), 0, a6, a12, 1.0, -1.0)
        assert n == 0


def test_oracle_excess_pairs_equal_launch_set_order():
    """orc_excess_pairs (MmSystem.cpp:700-736) gives the ala3 sample's excluded
    pairs in the order the host side hands to the engine; with them the self
    energy reproduces the sample log (test_zd_constants_and_self_energy_kat)."""
    ls, _ = ala3_launchset()
    tab = np.ascontiguousarray(ls.nb15off_table(), np.int32)
    cap = 4 * len(ls.nb15off) + 16
    out = np.zeros((cap, 2), np.int32)
    n = oracle.lib().orc_excess_pairs(ctypes.c_int(ls.n_atoms), tab.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.c_int(cap))
    assert n == len(ls.excess_pairs()) > 0
    assert np.array_equal(out[:n], ls.excess_pairs())


# ------------------------------ angles, torsions, impropers, 1-4 and SHAKE quad

def _raw_log_terms(d):
    """Angle, Torsion, Improper, 14-VDW, 14-Ele per step from a fixture's log lines."""
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


def test_ala3_bonded_terms_match_log():
    """The ala3 sample's own launch set at step 0: angle, torsion, improper,
    1-4 VDW and 1-4 Ele energies to the printed digits of the sample log
    (samples/ala3/log03_md.txt:131-133; the same values head the reference
    run in wons_ala3_traj.npz)."""
    d = load_npz("wons_ala3_traj.npz")
    ls = cls_from_fixture(d)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    e = oracle.bonded(ls, crd, np.zeros((ls.n_atoms, 3)))
    ref = _raw_log_terms(d)[0]
    assert (ref["angle"], ref["torsion"], ref["impro"]) == (4.1796133554, 28.320798304, 0.064242037627)
    for k in ("angle", "torsion", "impro", "vdw14", "ele14"):
        assert fmt(e[k]) == fmt(ref[k]), k


def test_oracle_ala3_gradient_matches_reference_one_step():
    """One leapfrog step of the ala3 launch set without constraints in the
    reference all-pairs build: the whole gradient (pairs, excluded pairs,
    bonds, angles, torsions, impropers, 1-4) to 1e-11 RMS."""
    from helpers import oracle_forces_allpairs
    d = load_npz("wons_ala3_grad.npz")
    ls = cls_from_fixture(d)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    work, evdw, eele, eb = oracle_forces_allpairs(ls, crd, oracle.zd_setup(12.0, 0.0), L)
    oracle.bonded(ls, crd, work)
    ref = d["work"]
    assert rms_rel(work, ref) < 1e-11
    assert np.max(np.abs(work - ref)) < 1e-8 * np.max(np.abs(ref))


def test_oracle_ala3_md_with_quads_matches_reference_run():
    """10 leapfrog steps of ala3 with SHAKE pairs, trios and four-atom groups
    and COM cancel on the solute and solvent groups: per-step bonded terms,
    potential and kinetic energy, and the final restart."""
    from helpers import oracle_md
    d = load_npz("wons_ala3_traj.npz")
    ls = cls_from_fixture(d)
    assert len(ls.shake_quads) == 9
    groups = [g for name, g in ls.groups if name in ("solute", "solvent")]
    crd, vel, rows = oracle_md(ls, 10, constraint="shake", com_groups=groups)
    log = d["log"]
    for r, lg in zip(rows, log[:10]):
        assert abs(r[1] - lg[1]) <= 2e-9 * abs(lg[5]) + 1e-8, (r, lg)
        assert abs(r[2] - lg[2]) <= 1e-9 * abs(lg[2]), (r, lg)
    assert np.max(np.abs(crd - d["crd"])) < 1e-9
    assert np.max(np.abs(vel - d["vel"])) < 1e-11


def test_shake_quad_restores_distances():
    """Closed check: after orc_shake_quad every one of the six distances of
    each group is at its target to the tolerance, and a group already at its
    targets does not move."""
    d = load_npz("wons_ala3_traj.npz")
    ls = cls_from_fixture(d)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    ids, d2 = ls.shake_quad_tables()
    minv = (1.0 / ls.mass.astype(np.float32).astype(np.float64)).astype(np.float32)
    rng = np.random.default_rng(5)
    prev = ls.crd.copy()
    cur = prev + rng.normal(scale=0.02, size=prev.shape)
    out, nf = oracle.shake_quad(L, ids, d2, cur, prev, minv, 1e-12, 1000)
    assert nf == 0
    pairs = [(0, 1), (1, 2), (2, 0), (3, 0), (2, 3), (1, 3)]
    for q in range(len(ids)):
        for p, (i, j) in enumerate(pairs):
            v = out[ids[q, i]] - out[ids[q, j]]
            assert abs(v @ v - d2[q, p]) <= 1e-10 * d2[q, p]
    out2, _ = oracle.shake_quad(L, ids, d2, out, out, minv, 1e-12, 1000)
    assert np.max(np.abs(out2 - out)) < 1e-12


# ------------------------------------- zero-quadrupole, -octupole, -hexadecapole

ZM_NAMES = {"q": "zero-quadrupole", "o": "zero-octupole", "h": "zero-hexadecapole"}
ZM_CASES = [(t, a) for t in "qoh" for a in "01"]


def _zm_poly(zd):
    """(core, {power: coefficient}) of the selected order's pair energy
    1/r (erfc) - core - sum c_k r^k."""
    m = zd.mode
    c = {0: {2: -zd.bcoeff}, 1: {2: zd.zq2, 4: zd.zq4}, 2: {2: zd.zo2, 4: zd.zo4, 6: zd.zo6},
         3: {2: zd.zh2, 4: zd.zh4, 6: zd.zh6, 8: zd.zh8}}[m]
    return zd.zmcore, c


@pytest.mark.parametrize("tag,atag", ZM_CASES)
def test_oracle_zero_multipole_gradient_matches_reference_one_step(tag, atag):
    """orc_zm_setup / the pair, excess and self terms of every higher order at
    alpha 0 and 0.1 against one step of the reference all-pairs build
    (wons_zm.npz): gradient to 1e-11 RMS, VDW and Ele to the printed digits."""
    d = load_npz("wons_zm.npz")
    ls = cls_from_fixture(d)
    alpha = 0.1 if atag == "1" else 0.0
    zd = oracle.zd_setup(12.0, alpha, oracle.ZM_MODES[ZM_NAMES[tag]])
    L, lb, ub = oracle.pbc_from_box(ls.box)
    crd = wrap_inbox(ls.crd, L, lb, ub)
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    work, en, _ = oracle.allpairs_fp32(zd, L, crd, ls.atom_type, lj6, lj12, q, ls.nb15off_table())
    ex = ls.excess_pairs()
    eb = oracle.bonds(L, ls.bond_ids, ls.bond_k, ls.bond_r0, crd, work)
    eex = oracle.excess(zd, L, ex, crd, q, work, bug_compat=True)
    eself, _ = oracle.self_energy(zd, q, ex)
    ref = d["work_" + tag + atag]
    assert rms_rel(work, ref) < 1e-11
    assert np.max(np.abs(work - ref)) < 1e-8 * np.max(np.abs(ref))
    log0 = d["log_" + tag + atag][0]
    assert fmt(en[0]) == fmt(log0[4])
    assert fmt(en[1] + eex + eself) == fmt(log0[5])
    assert abs(eb - log0[3]) <= 1e-9 * max(1.0, abs(log0[3]))


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("alpha", [0.0, 0.1])
def test_zero_multipole_constants_closed_form(mode, alpha):
    """The order-l zero-multipole pair energy erfc(ar)/r - core - sum c_k r^k
    and its first l derivatives vanish at the cutoff (l = 1 dipole ... 4
    hexadecapole), to the float rounding of the constants; the one exception
    is the hexadecapole's 4th derivative at alpha > 0, which carries the
    reference's series for the 4th derivative coefficient
    (ZeroMultipoleSum.cpp:152-154) and is checked to miss."""
    sp = pytest.importorskip("sympy")
    zd = oracle.zd_setup(12.0, alpha, mode)
    core, c = _zm_poly(zd)
    r = sp.symbols("r")
    f = sp.erfc(alpha * r) / r
    E = f - core - sum(v * r ** k for k, v in c.items())
    for n in range(mode + 2):
        val = float(sp.diff(E, r, n).subs(r, 12.0))
        scale = abs(float(sp.diff(f, r, n).subs(r, 12.0)))
        if mode == 3 and alpha > 0 and n == 4:
            assert abs(val) > 0.1 * scale
        else:
            assert abs(val) < 3e-6 * scale, (n, val, scale)
    # the next derivative does not vanish (the order is not higher than stated)
    n = mode + 2
    assert abs(float(sp.diff(E, r, n).subs(r, 12.0))) > 0.1 * abs(float(sp.diff(f, r, n).subs(r, 12.0)))


def _fd_check(efun, r, h=2e-2):
    """(g, dE/dr / r by central difference): the gradient coefficient of a
    pair term against its energy."""
    e1, _ = efun(r + h)
    e0, _ = efun(r - h)
    _, g = efun(r)
    return g, (e1 - e0) / (2 * h) / r


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("alpha", [0.0, 0.1])
def test_zero_multipole_pair_and_excess_forms_consistent(mode, alpha):
    """Pair term: the gradient coefficient equals dE/dr / r for every order
    (with the octupole alpha > 0 gradient on its own coefficients), and the
    reference's octupole alpha > 0 gradient, which takes the hexadecapole
    coefficients (ZeroMultipoleSum.cpp:405-406), does not.  Excluded-pair term
    in the fixed forms: consistent for every order; the reference forms miss
    where they drop or misplace a power (octupole r^5, hexadecapole)."""
    L = np.array([60.0, 60.0, 60.0], np.float32)
    q = np.array([0.8, -0.6], np.float32)

    def pair(zd):
        def e(r):
            n, ev, ee, w = oracle.pair_fp32(zd, np.array([r, 0.0, 0.0]), 0.0, 0.0, q[0], q[1])
            return ee, w[0] / r
        return e

    def excess(zd, bug):
        def e(r):
            crd = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]])
            w = np.zeros((2, 3))
            ene = oracle.excess(zd, L, np.array([[0, 1]], np.int32), crd, q, w, bug_compat=bug)
            # work is dE/dx_a = g * (x_a - x_b) = -g r along x
            return ene, -w[0, 0] / r
        return e

    for r in (4.0, 7.5, 10.5):
        zd = oracle.zd_setup(12.0, alpha, mode, zo_grad_bug=False)
        g, fd = _fd_check(pair(zd), r)
        # float32 energies: the difference quotient carries ~1e-6 / h of rounding
        assert g == pytest.approx(fd, rel=2e-3, abs=2e-5), (r, g, fd)
    if mode == 2 and alpha > 0:
        zd = oracle.zd_setup(12.0, alpha, mode, zo_grad_bug=True)
        g, fd = _fd_check(pair(zd), 7.5)
        assert abs(g - fd) > 1e-2 * abs(fd)
    for r in (1.0, 1.6, 2.5):
        zd = oracle.zd_setup(12.0, alpha, mode)
        g, fd = _fd_check(excess(zd, False), r)
        assert g == pytest.approx(fd, rel=2e-3, abs=2e-5), (r, g, fd)
    ref_slip = (mode == 2) or (mode == 3) or (mode == 0 and alpha > 0)
    if ref_slip:
        # at a separation where the misplaced powers weigh (excluded pairs sit at 1-3 A, where the octupole and
        # hexadecapole slips change the terms by less than 1e-4)
        g, fd = _fd_check(excess(oracle.zd_setup(12.0, alpha, mode), True), 9.0)
        assert abs(g - fd) > 1e-2 * abs(fd), (g, fd)


def test_oracle_zero_quadrupole_md_matches_reference_run():
    """10 leapfrog steps with SHAKE under zero-quadrupole at alpha 0: per-step
    potential and kinetic energy and the final restart (wons_zm.npz)."""
    from helpers import oracle_md
    d = load_npz("wons_zm.npz")
    from omegagene_amd.cls import read_cls
    ls = read_cls(bytes(d["traj_cls"]))
    zd = oracle.zd_setup(12.0, 0.0, oracle.ZM_MODES["zero-quadrupole"])
    crd, vel, rows = oracle_md(ls, 10, constraint="shake", zd=zd)
    log = d["traj_log"]
    for r, lg in zip(rows, log[:10]):
        assert abs(r[1] - lg[1]) <= 2e-9 * abs(lg[5]) + 1e-8, (r, lg)
        assert abs(r[2] - lg[2]) <= 1e-9 * abs(lg[2]), (r, lg)
    assert np.max(np.abs(crd - d["traj_crd"])) < 1e-9
    assert np.max(np.abs(vel - d["traj_vel"])) < 1e-11


# ------------------------------------------------------------ V-McMD, V-AUS

def _ext_setup(d, tag):
    from omegagene_amd.cls import read_cls
    ls = read_cls(bytes(d[tag + "_cls"]))
    ext = {"mode": "v-mcmd" if tag == "mcmd" else "v-aus", "vmcmd": ls.vmcmd}
    if tag != "mcmd":
        gc = ls.group_coord
        names = [nm for nm, _ in ls.groups]
        ga = np.asarray(ls.groups[gc.group_ids[0] - 1][1], np.int32)
        gb = np.asarray(ls.groups[gc.group_ids[1] - 1][1], np.int32)
        assert names[gc.group_ids[0] - 1] == "end_a"
        ext.update(groups=(ga, gb), crdg=np.ascontiguousarray(np.concatenate(gc.crd), np.float64), sigma=0.5,
                   recov=2.0)
    return ls, ext


@pytest.mark.parametrize("tag", ["mcmd", "aus", "aus_rec"])
def test_oracle_vmcmd_vaus_match_reference_runs(tag):
    """20 leapfrog steps with SHAKE under V-McMD (lambda = potential energy,
    the whole gradient scaled) and V-AUS (lambda = end-to-end mass-centre
    distance; state moves every trial, or a recovery force with no move)
    against the reference runs (wons_ext.npz): lambda per step to the printed
    11 digits, the virtual-state log, per-step energies and the restart."""
    from helpers import oracle_md
    d = load_npz("wons_ext.npz")
    ls, ext = _ext_setup(d, tag)
    crd, vel, rows = oracle_md(ls, 20, constraint="shake", ext=ext)
    # lambda of the reference's extra last step (step 20) too
    got = []
    ls, ext = _ext_setup(d, tag)
    oracle_md(ls, 21, constraint="shake", ext=ext, ext_out=got)
    lam_ref = d[tag + "_lambda"]
    assert len(lam_ref) == 20
    for k in range(1, 21):   # the reference writes lambda for steps 1..n (Extend.cpp:136)
        assert "%.10e" % got[k][0] == "%.10e" % lam_ref[k - 1], (k, got[k][0], lam_ref[k - 1])
    vs_ref = d[tag + "_vs"]
    for step, vs1 in vs_ref:
        assert got[step][1] == vs1 - 1, (step, got[step][1], vs1)
    n_states = len(set(v for _, v in got))
    assert n_states == 1 if tag == "aus_rec" else n_states >= 2
    log = d[tag + "_log"]
    for r, lg in zip(rows, log[:20]):
        assert abs(r[1] - lg[1]) <= 2e-9 * abs(lg[5]) + 1e-8, (r, lg)
        assert abs(r[2] - lg[2]) <= 1e-9 * abs(lg[2]), (r, lg)
    assert np.max(np.abs(crd - d[tag + "_crd"])) < 1e-9
    assert np.max(np.abs(vel - d[tag + "_vel"])) < 1e-11


def test_vs_transition_rule():
    """orc_vs_transition on its own: up is tried first, a move needs lambda in
    both closed ranges and the draw above 1 - p; the edges never move out."""
    lo, hi = [0.0, 4.0, 8.0], [5.0, 9.0, 12.0]
    pdn, pup = [0.0, 0.3, 1.0], [0.6, 0.7, 0.0]
    T = oracle.vs_transition
    assert T(lo, hi, pdn, pup, 0, 4.5, 0.5, 0.99) == 1        # 0.5 > 1 - 0.6
    assert T(lo, hi, pdn, pup, 0, 4.5, 0.3, 0.99) == 0        # 0.3 <= 0.4; down impossible at 0
    assert T(lo, hi, pdn, pup, 1, 4.5, 0.99, 0.8) == 0        # lambda not in state 2; down: 0.8 > 1 - 0.3
    assert T(lo, hi, pdn, pup, 1, 4.5, 0.99, 0.5) == 1        # 0.5 <= 0.7: stays
    assert T(lo, hi, pdn, pup, 1, 5.0, 0.99, 0.71) == 0       # closed range: 5.0 is in state 0
    assert T(lo, hi, pdn, pup, 1, 8.5, 0.31, 0.99) == 2       # up first
    assert T(lo, hi, pdn, pup, 1, 9.5, 0.99, 0.99) == 1       # lambda outside the source's range
    assert T(lo, hi, pdn, pup, 2, 8.5, 0.99, 0.5) == 1        # p_down 1
    assert T(lo, hi, pdn, pup, 2, 8.5, 0.99, 0.0) == 2        # the draw must exceed 1 - p
