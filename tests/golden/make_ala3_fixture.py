"""Write tests/golden/ala3_step0.npz from the reference sample samples/ala3.

Inputs: samples/ala3/md.inp.cls (the only v.0.52 launch set the reference
ships) and the step-0 energy terms printed in samples/ala3/log03_md.txt:99-135
(all-pairs build, 12 A, zero-dipole alpha 0).  The fixture keeps only the
arrays the pair/ZD path needs (coordinates, charges, types, LJ tables,
exclusions, box) plus the logged numbers; it is data, not reference code.

Run in a container where the reference tree is mounted:
    python tests/golden/make_ala3_fixture.py
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from omegagene_amd.cls import read_cls  # noqa: E402

REF = os.environ.get("OG_REFERENCE", "/root/reference")


def main():
    ls = read_cls(os.path.join(REF, "samples/ala3/md.inp.cls"))
    log = open(os.path.join(REF, "samples/ala3/log03_md.txt")).read()

    def grab(label):
        m = re.search(r"(?m)^\s*" + label + r"\s*:?\s*([-+0-9.eE]+)", log)
        return float(m.group(1))

    expected = dict(
        vdw=grab(r"VDW:"), ele=float(re.search(r"(?m)^VDW:\s*\S+\s+Ele:\s*(\S+)", log).group(1)), fcoeff=grab("fcoeff"), bcoeff=grab("bcoeff"),
        zmcore=grab("zmcore"), self_sum=grab("sum of energy self"), d_free=grab("Degree of freedom"),
        ke_init=grab("Initial kinetic energy"), t_init=grab("Initial temperature"))
    lj6, lj12 = ls.lj6, ls.lj12
    np.savez_compressed(
        os.path.join(HERE, "ala3_step0.npz"), box=ls.box, crd=ls.crd, vel=ls.vel, charge=ls.charge, mass=ls.mass,
        atom_type=ls.atom_type, lj6=lj6, lj12=lj12, nb15off=ls.nb15off,
        n_shake=np.array([len(ls.shake_pair_ids), len(ls.shake_trio_ids), len(ls.shake_quads)]),
        has_pos_rest=np.array(ls.has_pos_rest),
        expected_keys=np.array(list(expected.keys())), expected_vals=np.array(list(expected.values())))
    print("wrote ala3_step0.npz", expected)


if __name__ == "__main__":
    main()
