"""Write tests/golden/wons_*.npz from runs of the reference all-pairs build
(omegagene_wons, compiled from the reference sources by oracle/Makefile).

  wons_grad_a0.npz / wons_grad_a1.npz -- one leapfrog step without constraints
      on a water + 20-bead-chain system (ZD alpha 0 and 0.1).  The per-atom
      gradient follows from the restart velocities:
      work = -(v_half - v_in) / (dt * 4.184e-4 * mass_inv_f32)
      (SubBox.cpp:1445-1454; restart writer WriteTrr.cpp:258-313).
  wons_traj.npz -- 20 leapfrog steps of 1,000 rigid waters with SHAKE; the
      per-step energy terms of the log and the final restart.
  wons_langevin.npz -- position Langevin (gamma 10/ps, 300 K, seed 1234, no
      constraints: DynamicsMode.cpp:877) on water + a 20-bead chain, 3 steps
      plus the reference's extra last step.
  wons_shake_pair.npz -- 10 leapfrog steps with SHAKE pairs on every third
      bond of a 40-bead chain and SHAKE trios on the waters.
  wons_com.npz -- 10 leapfrog steps with SHAKE trios on the waters and
      --com-motion cancel on the atom groups "solute" and "solvent".
  wons_thermo.npz -- 10 leapfrog steps with --thermostat scaling (no
      constraints, so ThermostatScaling::apply_thermostat), --berendsen-tau
      0.1 ps, heating from 250 K to 300 K over 5 steps.

  wons_hpsdh_grad.npz -- one leapfrog step without constraints of 8 coarse-
      grained 60-bead chains (omegagene_amd.gen.build_cg_system) with
      --nonbond hydrophobicity-scale-lj and --electrostatic debye-huckel
      (cutoff 20 A, list 23 A, dielectric 85, 300 K, ionic strength 150,
      HPS epsilon 0.2; the cg_q8 sample's settings with a shorter Debye
      length); the gradient from the restart velocities as above.
  wons_hpsdh_traj.npz -- 20 leapfrog steps of dt 5 fs of the same form at the
      sample's ionic strength 0.00015; log and final restart.

  wons_ala3_grad.npz -- one leapfrog step without constraints of the ala3
      sample's own launch set (samples/ala3/md.inp.cls: angles, torsions,
      impropers, 1-4 pairs); the gradient from the restart velocities.
  wons_ala3_traj.npz -- 10 leapfrog steps of the same launch set with SHAKE
      (pairs, trios and the four-atom groups, tolerance 1e-10) and COM cancel
      on the solute and solvent groups, as the sample runs it; log and restart.
  wons_zm.npz -- one leapfrog step without constraints of the wons_grad
      system under --electrostatic zero-quadrupole, zero-octupole and
      zero-hexadecapole at --ele-alpha 0 and 0.1 (gradient from the restart
      velocities, log), and 10 leapfrog steps with SHAKE under
      zero-quadrupole at alpha 0 (log and final restart).
  wons_ext.npz -- 20 leapfrog steps with SHAKE on the waters of a water +
      20-bead-chain system under --extended-ensemble v-mcmd (three virtual
      states over the potential energy, transitions every 2 steps with
      probabilities 0 or 1 so the walk does not depend on the random stream)
      and the same under v-aus on the distance between the chain's end
      groups (group coordinates in the launch set); logs, lambda files,
      virtual-state logs and final restarts.
  wons_output.npz -- the reference's own output files of a 20-step leapfrog
      run with SHAKE (500 waters + 20-bead chain): the Presto trajectory
      (--print-interval-coord 5), the same run's GROMACS .trr, and the final
      restart, as raw bytes.

Inputs are the .cls bytes written by omegagene_amd.gen (stored in the fixture).
Run where the reference tree is mounted:
    make -C oracle ref && python tests/golden/make_wons_fixtures.py [more|hps|out|ala3|zm|ext]
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from omegagene_amd import gen  # noqa: E402
from omegagene_amd.cls import write_cls  # noqa: E402

sys.path.insert(0, os.path.dirname(HERE))
from helpers import chain_shake_pairs  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "omegagene_wons")


def read_restart(path):
    b = open(path, "rb").read()
    o = 4 + 80 + 4
    n = np.frombuffer(b, np.int32, 2, o + 4)[0]
    o += 4 + 8 + 4
    step = np.frombuffer(b, np.int32, 1, o + 4)[0]
    t, etot, ekin, epot = np.frombuffer(b, np.float64, 4, o + 8)
    o += 4 + 36 + 4
    crd = np.frombuffer(b, np.float64, 3 * n, o + 4).reshape(n, 3).copy()
    o += 4 + 24 * n + 4
    vel = np.frombuffer(b, np.float64, 3 * n, o + 4).reshape(n, 3).copy()
    return dict(n=int(n), step=int(step), time=t, etot=etot, ekin=ekin, epot=epot, crd=crd, vel=vel)


def parse_log(text):
    rows = []
    for blk in re.split(r"(?m)^Step:", text)[1:]:
        def g(lbl):
            m = re.search(lbl + r"\s*([-+0-9.eE]+)", blk)
            return float(m.group(1)) if m else np.nan
        rows.append([float(blk.split()[0]), g(r"(?m)^Potential:"), g("Kinetic:"), g(r"(?m)^Bond:"), g(r"(?m)^VDW:"),
                     g(r"(?m)^VDW:\s*\S+\s+Ele:"), g(r"(?m)^Temperature:")])
    return np.array(rows)


def run(ls, cfg, workdir, extra=""):
    clsf = os.path.join(workdir, "in.cls")
    cfgf = os.path.join(workdir, "in.cfg")
    if isinstance(ls, bytes):  # a launch set file as it is
        data = ls
        open(clsf, "wb").write(data)
    else:
        data = write_cls(ls, clsf)
    with open(cfgf, "w") as f:
        f.write(gen.cfg_text(cfg) + extra)
    r = subprocess.run([EXE, "--cfg", cfgf, "--inp", clsf], cwd=workdir, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    # the reference may fail in its destructors after finishing the run; the
    # run itself is complete once terminal_process has been printed
    if "terminal_process" not in r.stdout:
        raise RuntimeError(r.stdout[-2000:])
    return data, r.stdout, read_restart(os.path.join(workdir, "out.restart"))


BASE = {"mode": "md", "integrator": "leapfrog-presto", "cutoff": 12.0, "nsgrid-cutoff": 13.0,
        "nsgrid-update-intvl": 1, "time-step": 2.0, "electrostatic": "zero-dipole", "thermostat": "none",
        "com-motion": "none", "print-interval-log": 1, "fn-o-restart": "out.restart", "random-seed": 1}


def more(tmp):
    out = {}
    ls = gen.build_system(500, seed=41, n_beads=20)
    cfg = dict(BASE, **{"integrator": "langevin", "langevin-gamma": 10.0, "temperature": 300.0, "random-seed": 1234,
                        "n-steps": 3, "ele-alpha": 0.0, "constraint": "none"})
    out["wons_langevin"] = (ls, cfg, "")
    ls = chain_shake_pairs(gen.build_system(500, seed=43, n_beads=40))
    cfg = dict(BASE, **{"n-steps": 10, "ele-alpha": 0.0, "constraint": "shake", "const-tolerance": 1e-10,
                        "const-max-loops": 1000})
    out["wons_shake_pair"] = (ls, cfg, "")
    ls = gen.build_system(500, seed=45, n_beads=20)
    cfg = dict(BASE, **{"n-steps": 10, "ele-alpha": 0.0, "constraint": "shake", "const-tolerance": 1e-10,
                        "const-max-loops": 1000, "com-motion": "cancel"})
    out["wons_com"] = (ls, cfg, "--com-cancel-group-name      solute\n--com-cancel-group-name      solvent\n")
    ls = gen.build_system(500, seed=47, n_beads=20)
    cfg = dict(BASE, **{"n-steps": 10, "ele-alpha": 0.0, "constraint": "none", "thermostat": "scaling",
                        "berendsen-tau": 0.1, "temperature": 300.0, "temperature-init": 250.0, "heating-steps": 5})
    out["wons_thermo"] = (ls, cfg, "")
    for name, (ls, cfg, extra) in out.items():
        data, log, rs = run(ls, cfg, tmp, extra)
        lg = parse_log(log)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), cls=np.frombuffer(data, np.uint8), log=lg,
                            crd=rs["crd"], vel=rs["vel"], epot=rs["epot"], ekin=rs["ekin"], step=rs["step"],
                            cfg=np.frombuffer((gen.cfg_text(cfg) + extra).encode(), np.uint8))
        print(name, "ok", lg.shape, lg[0], lg[-1])


HPS_DH = {"cutoff": 20.0, "nsgrid-cutoff": 23.0, "electrostatic": "debye-huckel", "debye-huckel-dielectric": 85.0,
          "debye-huckel-temperature": 300.0, "nonbond": "hydrophobicity-scale-lj",
          "hydrophobicity-scale-epsiron": 0.2, "constraint": "none", "ele-alpha": 0.0}


def hps(tmp):
    ls = gen.build_cg_system(8, 60, seed=3, L=90.0)
    cfg = dict(BASE, **HPS_DH, **{"n-steps": 1, "time-step": 2.0, "debye-huckel-ionic-strength": 150.0})
    data, log, rs = run(ls, cfg, tmp)
    minv = (1.0 / ls.mass.astype(np.float32).astype(np.float64)).astype(np.float32).astype(np.float64)
    work = -(rs["vel"] - ls.vel) / (2.0 * 4.184e-4 * minv[:, None])
    np.savez_compressed(os.path.join(HERE, "wons_hpsdh_grad.npz"), cls=np.frombuffer(data, np.uint8), work=work,
                        log=parse_log(log), crd1=rs["crd"], vel1=rs["vel"],
                        cfg=np.frombuffer(gen.cfg_text(cfg).encode(), np.uint8))
    print("hpsdh grad ok", parse_log(log)[0])
    cfg = dict(BASE, **HPS_DH, **{"n-steps": 20, "time-step": 5.0, "debye-huckel-ionic-strength": 0.00015})
    data, log, rs = run(ls, cfg, tmp)
    lg = parse_log(log)
    np.savez_compressed(os.path.join(HERE, "wons_hpsdh_traj.npz"), cls=np.frombuffer(data, np.uint8), log=lg,
                        crd=rs["crd"], vel=rs["vel"], epot=rs["epot"], ekin=rs["ekin"], step=rs["step"],
                        cfg=np.frombuffer(gen.cfg_text(cfg).encode(), np.uint8))
    print("hpsdh traj ok", lg.shape, lg[0], lg[-1])


def outputs(tmp):
    ls = gen.build_system(500, seed=49, n_beads=20)
    cfg = dict(BASE, **{"n-steps": 20, "ele-alpha": 0.0, "constraint": "shake", "const-tolerance": 1e-10,
                        "const-max-loops": 1000, "print-interval-coord": 5})
    files = {}
    for fmt in ("presto", "gromacs"):
        c = dict(cfg, **{"fn-o-coord": "out.trr", "format-o-coord": fmt})
        data, log, rs = run(ls, c, tmp)
        files[fmt] = open(os.path.join(tmp, "out.trr"), "rb").read()
        files["restart"] = open(os.path.join(tmp, "out.restart"), "rb").read()
    np.savez_compressed(os.path.join(HERE, "wons_output.npz"), cls=np.frombuffer(data, np.uint8),
                        log=parse_log(log), trr_presto=np.frombuffer(files["presto"], np.uint8),
                        trr_gromacs=np.frombuffer(files["gromacs"], np.uint8),
                        restart=np.frombuffer(files["restart"], np.uint8),
                        cfg=np.frombuffer(gen.cfg_text(dict(cfg, **{"format-o-coord": "presto"})).encode(), np.uint8))
    print("outputs ok", len(files["presto"]), len(files["gromacs"]), len(files["restart"]))


def ala3(tmp):
    data = open("/root/reference/samples/ala3/md.inp.cls", "rb").read()
    base = dict(BASE, **{"ele-alpha": 0.0, "const-tolerance": 1e-10, "const-max-loops": 1000})
    cfg = dict(base, **{"n-steps": 1, "constraint": "none"})
    from omegagene_amd.cls import read_cls
    ls = read_cls(data)
    _, log, rs = run(data, cfg, tmp)
    minv = (1.0 / ls.mass.astype(np.float32).astype(np.float64)).astype(np.float32).astype(np.float64)
    work = -(rs["vel"] - ls.vel) / (2.0 * 4.184e-4 * minv[:, None])
    np.savez_compressed(os.path.join(HERE, "wons_ala3_grad.npz"), cls=np.frombuffer(data, np.uint8), work=work,
                        log=parse_log(log), crd1=rs["crd"], vel1=rs["vel"],
                        cfg=np.frombuffer(gen.cfg_text(cfg).encode(), np.uint8))
    print("ala3 grad ok", parse_log(log)[0])
    extra = "--com-cancel-group-name      solute\n--com-cancel-group-name      solvent\n"
    cfg = dict(base, **{"n-steps": 10, "constraint": "shake", "com-motion": "cancel"})
    _, log, rs = run(data, cfg, tmp, extra)
    lg = parse_log(log)
    np.savez_compressed(os.path.join(HERE, "wons_ala3_traj.npz"), cls=np.frombuffer(data, np.uint8), log=lg,
                        crd=rs["crd"], vel=rs["vel"], epot=rs["epot"], ekin=rs["ekin"], step=rs["step"],
                        cfg=np.frombuffer((gen.cfg_text(cfg) + extra).encode(), np.uint8), raw_log=np.frombuffer(
                            "\n".join(l for l in log.splitlines() if l[:5] in ("Step:", "Total", "Poten", "Bond:",
                                                                               "Torsi", "14-VD", "VDW: ", "Tempe")
                                       ).encode(), np.uint8))
    print("ala3 traj ok", lg.shape, lg[0], lg[-1])


ZM = {"q": "zero-quadrupole", "o": "zero-octupole", "h": "zero-hexadecapole"}


def zm(tmp):
    ls = gen.build_system(700, seed=77, n_beads=20)
    minv = (1.0 / ls.mass.astype(np.float32).astype(np.float64)).astype(np.float32).astype(np.float64)
    out = {}
    for tag, ele in ZM.items():
        for atag, alpha in (("0", 0.0), ("1", 0.1)):
            cfg = dict(BASE, **{"n-steps": 1, "ele-alpha": alpha, "constraint": "none", "electrostatic": ele})
            data, log, rs = run(ls, cfg, tmp)
            out["work_%s%s" % (tag, atag)] = -(rs["vel"] - ls.vel) / (2.0 * 4.184e-4 * minv[:, None])
            out["log_%s%s" % (tag, atag)] = parse_log(log)
            print(ele, alpha, parse_log(log)[0])
    tl = gen.build_system(1000, seed=3, n_beads=0)
    cfg = dict(BASE, **{"n-steps": 10, "ele-alpha": 0.0, "constraint": "shake", "const-tolerance": 1e-10,
                        "const-max-loops": 1000, "electrostatic": "zero-quadrupole"})
    tdata, log, rs = run(tl, cfg, tmp)
    lg = parse_log(log)
    print("quadrupole traj", lg.shape, lg[0], lg[-1])
    np.savez_compressed(os.path.join(HERE, "wons_zm.npz"), cls=np.frombuffer(data, np.uint8), traj_cls=np.frombuffer(
        tdata, np.uint8), traj_log=lg, traj_crd=rs["crd"], traj_vel=rs["vel"],
        traj_cfg=np.frombuffer(gen.cfg_text(cfg).encode(), np.uint8), **out)


def _read_text(path):
    return open(path).read() if os.path.exists(path) else ""


def ext(tmp):
    from omegagene_amd.cls import GroupCoordDef, VmcmdDef
    sys.path.insert(0, os.path.dirname(HERE))
    import oracle
    from helpers import oracle_md
    out = {}
    files = {"fn-o-vmcmd-log": "vs.log", "fn-o-extended-lambda": "lambda.out", "format-o-extended-lambda": "ascii",
             "print-interval-extended-lambda": 1}
    base = dict(BASE, **{"n-steps": 20, "ele-alpha": 0.0, "constraint": "shake", "const-tolerance": 1e-10,
                         "const-max-loops": 1000, "temperature": 300.0})
    # V-McMD: lambda = potential energy; states over the unbiased run's range
    ls = gen.build_system(500, seed=51, n_beads=20)
    _, _, rows = oracle_md(ls, 20, constraint="shake")
    e_lo, e_hi = float(rows[:, 1].min()), float(rows[:, 1].max())
    span = e_hi - e_lo
    kT = (8.31451 / 4.184) * 1e-3 * 300.0
    cuts = [e_lo - 0.2 * span, e_lo + 0.35 * span, e_lo + 0.65 * span, e_hi + 0.2 * span]
    states = []
    for k in range(3):
        lo, hi = cuts[k] - 0.1 * span, cuts[k + 1] + 0.1 * span
        c1 = 2e-5 * (k + 1)
        c0 = (1.0 + 0.01 * k) / kT - c1 * 0.5 * (lo + hi)
        poly = [c0, c1] + ([1e-10] if k == 1 else [])
        states.append((len(poly) - 1, lo, hi, 0.0 if k == 0 else 1.0, 0.0 if k == 2 else 1.0, poly, 0.0, 0.0))
    init = int(np.searchsorted(cuts, rows[0, 1]) - 1)
    ls.vmcmd = VmcmdDef(2, 300.0, states, min(max(init, 0), 2), 5)
    cfg = dict(base, **{"extended-ensemble": "v-mcmd"}, **files)
    data, log, rs = run(ls, cfg, tmp)
    out["mcmd"] = (data, log, rs, cfg, _read_text(os.path.join(tmp, "lambda.out")),
                   _read_text(os.path.join(tmp, "vs.log")))
    # V-AUS: lambda = distance between the mass centres of the chain's end groups
    ls = gen.build_system(500, seed=53, n_beads=20, with_vcmd=True)
    ls.vcmd = None
    names = [nm for nm, _ in ls.groups]
    ga, gb = dict(ls.groups)["end_a"], dict(ls.groups)["end_b"]
    m = ls.mass
    lam0 = float(np.linalg.norm((ls.crd[ga] * m[ga, None]).sum(0) / m[ga].sum() -
                                (ls.crd[gb] * m[gb, None]).sum(0) / m[gb].sum()))
    ls.group_coord = GroupCoordDef("V-AUS restart", 0, [names.index("end_a") + 1, names.index("end_b") + 1],
                                   [ls.crd[ga].copy(), ls.crd[gb].copy()])
    cfg = dict(base, **{"extended-ensemble": "v-aus", "enhance-sigma": 0.5, "enhance-recovery-coef": 2.0,
                        "fn-o-aus-restart": "aus.restart"}, **files)
    # "aus": lambda starts in the overlap of states 1 and 2 (moves every trial);
    # "aus_rec": state 0 ends 1.5 A below lambda (recovery force, no move)
    for tag, cuts, init in (("aus", [lam0 - 8.0, lam0 - 3.0, lam0 - 0.2, lam0 + 6.0], 1),
                            ("aus_rec", [lam0 - 9.0, lam0 - 2.1, lam0 + 2.0, lam0 + 8.0], 0)):
        states = []
        for k in range(3):
            states.append((1, cuts[k] - 0.6, cuts[k + 1] + 0.6, 0.0 if k == 0 else 1.0, 0.0 if k == 2 else 1.0,
                           [0.4 - 0.3 * k, -0.05 * (k + 1)], 0.0, 0.0))
        ls.vmcmd = VmcmdDef(2, 300.0, states, init, 5)
        data, log, rs = run(ls, cfg, tmp)
        out[tag] = (data, log, rs, cfg, _read_text(os.path.join(tmp, "lambda.out")),
                    _read_text(os.path.join(tmp, "vs.log")))
    arrs = {}
    for tag, (data, log, rs, cfg, lam_txt, vs_txt) in out.items():
        lg = parse_log(log)
        lam = np.array([float(x) for x in lam_txt.split()])
        vs = np.array([[int(x) for x in ln.split()] for ln in vs_txt.splitlines() if ln.strip()], np.int64)
        arrs.update({tag + "_cls": np.frombuffer(data, np.uint8), tag + "_log": lg, tag + "_crd": rs["crd"],
                     tag + "_vel": rs["vel"], tag + "_lambda": lam, tag + "_vs": vs.reshape(-1, 2),
                     tag + "_cfg": np.frombuffer(gen.cfg_text(cfg).encode(), np.uint8)})
        print(tag, lg.shape, lg[0], lg[-1], "lambda", lam[:3], "vs", vs[:5].tolist())
    np.savez_compressed(os.path.join(HERE, "wons_ext.npz"), **arrs)


def main():
    tmp = tempfile.mkdtemp()
    if len(sys.argv) > 1 and sys.argv[1] == "ext":
        ext(tmp)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "zm":
        zm(tmp)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "ala3":
        ala3(tmp)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "out":
        outputs(tmp)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "more":
        more(tmp)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "hps":
        hps(tmp)
        return
    for tag, alpha in (("a0", 0.0), ("a1", 0.1)):
        ls = gen.build_system(700, seed=77, n_beads=20)
        cfg = dict(BASE, **{"n-steps": 1, "ele-alpha": alpha, "constraint": "none"})
        data, log, rs = run(ls, cfg, tmp)
        minv = (1.0 / ls.mass.astype(np.float32).astype(np.float64)).astype(np.float32).astype(np.float64)
        work = -(rs["vel"] - ls.vel) / (2.0 * 4.184e-4 * minv[:, None])
        np.savez_compressed(os.path.join(HERE, "wons_grad_%s.npz" % tag), cls=np.frombuffer(data, np.uint8),
                            alpha=np.float64(alpha), work=work, log=parse_log(log), crd1=rs["crd"], vel1=rs["vel"])
        print(tag, "ok", parse_log(log)[0])
    ls = gen.build_system(1000, seed=88)
    cfg = dict(BASE, **{"n-steps": 20, "ele-alpha": 0.0, "constraint": "shake", "const-tolerance": 1e-10,
                        "const-max-loops": 1000})
    data, log, rs = run(ls, cfg, tmp)
    lg = parse_log(log)
    np.savez_compressed(os.path.join(HERE, "wons_traj.npz"), cls=np.frombuffer(data, np.uint8), log=lg,
                        crd=rs["crd"], vel=rs["vel"], epot=rs["epot"], ekin=rs["ekin"], step=rs["step"])
    print("traj ok", lg.shape, lg[0], lg[-1])


if __name__ == "__main__":
    main()
