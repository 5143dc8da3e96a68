"""Run outputs of the device engine (omegagene_amd.output.run_md over
og_io.cpp's writers) against the reference all-pairs build's own files for
the same 20-step leapfrog + SHAKE run (tests/golden/wons_output.npz): the
same frames at the same steps with the same layout, positions within 1e-4 A
(stored as float), energies within the FP32 pair-sum tolerance, and the
restart's step, time, positions and velocities."""
import sys

import numpy as np
import pytest

from helpers import cls_from_fixture, load_npz
from omegagene_amd.engine import System
from omegagene_amd.output import run_md
from test_output import gromacs_frames, presto_frames, restart_fields

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _cfg(d, **kw):
    c = {}
    for line in bytes(d["cfg"]).decode().splitlines():
        k, v = line.split()[:2]
        k = k[2:]
        try:
            v = int(v)
        except ValueError:
            try:
                v = float(v)
            except ValueError:
                pass
        c[k] = v
    c.update({"log-cap": 64})
    c.update(kw)
    return c


def test_run_outputs_match_reference_files(tmp_path):
    d = load_npz("wons_output.npz")
    ls = cls_from_fixture(d)
    say("run outputs", ls.n_atoms)
    ref_p = presto_frames(bytes(d["trr_presto"]))
    ref_g = gromacs_frames(bytes(d["trr_gromacs"]))
    ref_r = restart_fields(bytes(d["restart"]))
    for fmt in ("presto", "gromacs"):
        cfg = _cfg(d, **{"format-o-coord": fmt, "fn-o-coord": str(tmp_path / ("t_" + fmt)),
                         "fn-o-restart": str(tmp_path / "r.restart")})
        s = System(ls, cfg)
        try:
            run_md(s, ls, cfg, int(cfg["n-steps"]))
        finally:
            s.close()
        b = open(str(tmp_path / ("t_" + fmt)), "rb").read()
        if fmt == "presto":
            mine = presto_frames(b)
            assert [f["step"] for f in mine] == [f["step"] for f in ref_p]
            for a, r in zip(mine, ref_p):
                assert (a["m0"], a["m1"]) == (r["m0"], r["m1"])
                assert a["vals"][0] == r["vals"][0]                                  # time
                for k in (2, 3, 4, 5, 7):                                            # total, kin, T, pot, vdw
                    assert abs(a["vals"][k] - r["vals"][k]) <= 1e-5 * abs(r["vals"][k]) + 0.01, (a["step"], k)
                dx = np.max(np.abs(a["crd"] - r["crd"]))
                assert dx < 1e-4, (a["step"], dx)
            say("  presto: %d frames, max |dx| %.2e" % (len(mine), dx))
        else:
            mine = gromacs_frames(b)
            assert [f["step"] for f in mine] == [f["step"] for f in ref_g]
            hdr = 4 * 3 + 12 + 4 * 13 + 16 + 72
            for a, r in zip(mine, ref_g):
                assert a["raw"][:hdr] == r["raw"][:hdr]
                assert np.max(np.abs(a["crd"] - r["crd"])) < 1e-5
            say("  gromacs: %d frames" % len(mine))
    r = restart_fields(open(str(tmp_path / "r.restart"), "rb").read())
    assert (r["n"], r["step"], r["time"]) == (ref_r["n"], ref_r["step"], ref_r["time"])
    say("  restart: max |dx| %.2e, max |dv| %.2e, ekin %.6f / %.6f, epot %.4f / %.4f" %
        (np.max(np.abs(r["crd"] - ref_r["crd"])), np.max(np.abs(r["vel"] - ref_r["vel"])), r["ekin"], ref_r["ekin"],
         r["epot"], ref_r["epot"]))
    # FP32 pair sums diverge over 20 steps by ~1e-5 A on the flexible chain
    assert np.max(np.abs(r["crd"] - ref_r["crd"])) < 1e-4
    assert np.max(np.abs(r["vel"] - ref_r["vel"])) < 1e-5
    assert r["ekin"] == pytest.approx(ref_r["ekin"], rel=1e-6)
    assert abs(r["epot"] - ref_r["epot"]) <= 1e-5 * abs(ref_r["epot"]) + 0.01
