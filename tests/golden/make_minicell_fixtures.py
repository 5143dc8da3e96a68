"""Write tests/golden/minicell_*.npz: grid and cell-pair lists produced by the
reference MiniCell (src/celeste/core/MiniCell.cpp compiled from the reference
tree into oracle/_ref/minicell_drv by oracle/Makefile) on seeded systems.

Each fixture holds the inputs (box, coordinates, exclusion table, list
radius) and the reference outputs (atomids in grid order, grid coordinates,
column heads, cell-pair records and per-cell heads).  Run where the
reference tree is mounted:
    make -C oracle ref && python tests/golden/make_minicell_fixtures.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from omegagene_amd import gen  # noqa: E402

CASES = [
    # name, n_waters, seed, box, n_beads, list radius
    ("w1500", 1500, 5, None, 0, 13.0),
    ("cfg0", 2000, 1000, 39.1, 0, 13.0),
    ("chain", 1400, 7, None, 30, 13.0),
    ("chain200", 3000, 11, None, 200, 13.0),
    ("rc11", 2500, 21, None, 0, 11.0),
]


def main():
    tmp = tempfile.mkdtemp()
    for name, nw, seed, L, nb, rcl in CASES:
        ls = gen.build_system(nw, seed=seed, L=L, n_beads=nb)
        tab = ls.nb15off_table()
        r = oracle.run_ref_minicell(ls.n_atoms, ls.box, rcl, ls.crd, tab, tmp)
        np.savez_compressed(os.path.join(HERE, "minicell_%s.npz" % name), box=ls.box, crd=ls.crd, nb15off=tab,
                            rcl=np.float64(rcl), atomids=r["atomids"], grid_crd=r["crd"],
                            idx_xy_head_cell=r["idx_xy_head_cell"], idx_head_cell_pairs=r["idx_head_cell_pairs"],
                            cell_pairs=r["cell_pairs"].view(np.int32).reshape(-1, 5))
        print(name, ls.n_atoms, r["n_cells"], len(r["cell_pairs"]))


if __name__ == "__main__":
    main()
