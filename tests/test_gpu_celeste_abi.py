"""The reference's GPU call points (include/og_celeste.h), driven in the
order SubBox uses them (SubBox.cpp:487-558,1683-1729), against the oracle.

The host grid (atomids, grid-order coordinates, column heads) comes from the
oracle's MiniCell restatement, as MiniCell supplies it in the reference.
The device cell-pair list must equal the CPU list bit for bit; the grid-order
gradient must match the FP32-faithful oracle within RMS relative 2e-6 and
vdw within 2e-6 relative and ele within 1e-8 of the sum of the magnitudes
of its terms (helpers.ele_scale)."""
import sys

import numpy as np
import pytest

import oracle
from helpers import ele_scale, rms_rel, wrap_inbox
from omegagene_amd import gen
from omegagene_amd.celeste import SubBoxDriver

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _host_grid(o, c):
    g = o.g
    nx = g.n_cells_xyz[0]
    cell_col = c.cell_crd[:, 0] + c.cell_crd[:, 1] * nx
    return {"n_cells": c.n_cells, "atomids": c.atomids, "crd": c.crd, "idx_xy_head_cell": c.idx_xy_head_cell,
            "column_of_slot": np.repeat(cell_col, 8).astype(np.int32)}


def _geom(o):
    g = o.g
    return {"n_cells_xyz": [g.n_cells_xyz[0], g.n_cells_xyz[1], g.n_cells_xyz[2]],
            "L_cell_xyz": [g.L_cell_xyz[0], g.L_cell_xyz[1], g.L_cell_xyz[2]],
            "n_neighbors_xyz": [g.n_neighbors_xyz[0], g.n_neighbors_xyz[1], g.n_neighbors_xyz[2]],
            "n_columns": g.n_columns, "max_n_cells": g.max_n_cells, "max_n_cell_pairs": g.max_n_cell_pairs}


@pytest.mark.parametrize("case", ["cfg0", "chain"])
def test_reference_call_points_match_oracle(case):
    if case == "cfg0":
        ls, _ = gen.config0()
    else:
        ls = gen.build_system(1400, seed=7, n_beads=30)
    n = ls.n_atoms
    say("celeste call points", case, n)
    zd = oracle.zd_setup(12.0, 0.0)
    L, lb, ub = oracle.pbc_from_box(ls.box)
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    tab = ls.nb15off_table()
    o = oracle.Oracle(n, ls.box, 13.0)
    drv = None
    try:
        rng = np.random.default_rng(3)
        crd = ls.crd
        for rebuild in range(2):
            c = o.build(crd)
            grid = _host_grid(o, c)
            if drv is None:
                drv = SubBoxDriver(n, L, lb, 12.0, 13.0, zd, lj6, lj12, tab, q, ls.atom_type, _geom(o),
                                   max_n_atom_array=n + 8 * o.g.n_columns)
            drv.update_grid(grid)
            cps_ref, head_ref = o.enumerate(tab, gap_bug=True)
            cps, head = drv.gpu.cell_pairs(c.n_cells)
            assert np.array_equal(head, head_ref)
            assert np.array_equal(cps, cps_ref.view(np.int32).reshape(-1, 5))
            w, e = drv.pair_forces()
            wo, eo, npair = o.pairs_fp32(zd, cps_ref, ls.atom_type, lj6, lj12, q)
            m = c.atomids >= 0
            err = rms_rel(w[m], wo[m])
            say("  rebuild %d: %d cell pairs, rms rel %.3e, vdw %.6f/%.6f ele %.6f/%.6f" %
                (rebuild, len(cps), err, e[0], eo[0], e[1], eo[1]))
            assert err < 2e-6
            assert np.all(w[~m] == 0.0)
            assert e[0] == pytest.approx(eo[0], rel=2e-6, abs=1e-4)
            # ele is a near-cancelling sum: within 1e-8 of the sum of its term magnitudes
            assert abs(e[1] - eo[1]) <= 1e-8 * ele_scale(ls, crd=crd)
            # a second evaluation of the same coordinates is bit-identical (fixed point)
            w2, e2 = drv.pair_forces()
            assert np.array_equal(w, w2) and np.array_equal(e, e2)
            crd = wrap_inbox(crd + rng.normal(scale=0.05, size=crd.shape), L, lb, ub)
    finally:
        if drv is not None:
            drv.close()
