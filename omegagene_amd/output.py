"""The reference's run outputs on the device engine: trajectory frames and the
restart file, written by og_io.cpp's writers at the steps the reference
writes them.

``run_md`` plays DynamicsMode::main_stream (DynamicsMode.cpp:184-207) for one
``System``: before step k (k > 0, k a multiple of a --print-interval-coord)
a trajectory frame of the positions before the last update (copy_crd_prev)
with the energies of step k - 1 (sub_output, :272-300); after step k, k a
multiple of --print-interval-restart, the restart (output_restart,
:210-226); after the n steps a last frame, the restart, one more step and,
for Langevin, the restart again.  The engine runs the steps between two
outputs as one launch sequence.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .engine import _check, lib, potential

KINETIC_TO_K = 2.0 * 4.184 * 1e3 / 8.31451  # temperature_coeff * d_free (DynamicsMode.cpp:67)
TITLE = "omegagene_amd (MI355X)"


def _d(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def write_restart(path, n_steps, time_, e_potential, e_kinetic, crd, vel, title=TITLE):
    """og_write_restart (WriteRestart::write_restart, WriteTrr.cpp:258-313)."""
    c, pc = _d(crd)
    v, pv = _d(vel)
    _check(lib().og_write_restart(path.encode(), title.encode(), C.c_int(len(c)), C.c_int(int(n_steps)),
                                  C.c_double(time_), C.c_double(e_potential), C.c_double(e_kinetic), pc, pv))


def write_presto_frame(path, append, step, time_, cpu_time, total_e, kinetic_e, temperature, potential_e, vdw_e,
                       crd, group=None):
    """og_write_trr_presto_frame (WriteTrrPresto::write_trr, WriteTrr.cpp:179-254)."""
    c, pc = _d(crd)
    g = np.ascontiguousarray(group if group is not None else np.zeros(0), np.int32)
    _check(lib().og_write_trr_presto_frame(path.encode(), C.c_int(int(append)), C.c_int(len(c)), C.c_int(int(step)),
                                           C.c_float(time_), C.c_float(cpu_time), C.c_float(total_e),
                                           C.c_float(kinetic_e), C.c_float(temperature), C.c_float(potential_e),
                                           C.c_float(vdw_e), pc, C.c_int(len(g)),
                                           g.ctypes.data_as(C.POINTER(C.c_int))))


def write_gromacs_frame(path, append, step, time_, box_l, crd=None, vel=None, force=None, group=None):
    """og_write_trr_gromacs_frame (WriteTrrGromacs::write_trr, WriteTrr.cpp:33-172)."""
    arrs = [(_d(a) if a is not None else (None, None)) for a in (crd, vel, force)]
    n = next(len(a) for a, _ in arrs if a is not None)
    g = np.ascontiguousarray(group if group is not None else np.zeros(0), np.int32)
    _check(lib().og_write_trr_gromacs_frame(path.encode(), C.c_int(int(append)), C.c_int(n), C.c_int(int(step)),
                                            C.c_double(time_), C.c_double(box_l[0]), C.c_double(box_l[1]),
                                            C.c_double(box_l[2]), arrs[0][1], arrs[1][1], arrs[2][1],
                                            C.c_int(len(g)), g.ctypes.data_as(C.POINTER(C.c_int))))


def _crd_prev(system):
    out = np.zeros((system.n_atoms, 3))
    _check(lib().og_get_crd_prev(system.h, out.ctypes.data_as(C.c_void_p)))
    return out


def run_md(system, ls, cfg: dict, n_steps: int):
    """Run n_steps (plus the reference's extra last step) with the outputs of
    the .cfg keys print-interval-coord / fn-o-coord / format-o-coord /
    group-o-coord and print-interval-restart / fn-o-restart.  The system's
    energy log must keep every step (log-stride 1)."""
    as_list = lambda v: v if isinstance(v, (list, tuple)) else ([] if v is None else [v])  # noqa: E731
    intvls = [int(x) for x in as_list(cfg.get("print-interval-coord"))]
    fns = as_list(cfg.get("fn-o-coord"))
    fmt = cfg.get("format-o-coord", "presto")
    groups = as_list(cfg.get("group-o-coord"))
    rs_intvl = int(cfg.get("print-interval-restart", -1))
    rs_fn = cfg.get("fn-o-restart")
    dt = float(cfg.get("time-step", 2.0))
    langevin = cfg.get("integrator", "leapfrog-presto") == "langevin"
    names = {name: ids for name, ids in ls.groups}
    grp = [np.asarray(names[g], np.int32) if g in names else None for g in groups]
    started = [False] * len(fns)
    # PBC::L is real_pw (PBC.h:21): the box lengths rounded to float
    box_l = tuple(float(np.float32(x)) for x in (ls.box[0], ls.box[4], ls.box[8]))
    t_step = [0.0]

    def last_energies():
        row = system.energy_log(max_rows=1)[-1]
        return float(potential(row)), row[5], row[0], row[7]

    def frame(k):
        crd = _crd_prev(system)
        pot, kin, vdw, _ = last_energies()
        for i, (iv, fn) in enumerate(zip(intvls, fns)):
            if iv < 0 or k == 0 or k % iv != 0:
                continue
            g = grp[i] if i < len(grp) else None
            if fmt == "gromacs":
                write_gromacs_frame(fn, started[i], k, k * dt, box_l, crd=crd, group=g)
            else:
                write_presto_frame(fn, started[i], k, k * dt, t_step[0], pot + kin, kin,
                                   kin * KINETIC_TO_K / system.d_free, pot, vdw, crd, group=g)
            started[i] = True

    def restart(k):
        # cur_step and cur_time of the reference's loop at that point (the time advances after the output)
        if not rs_fn:
            return
        pot, kin, _, ext = last_energies()
        crd, vn, vj = system.state()
        if langevin:
            crd, vel = _crd_prev(system), vj
        else:
            vel = vn
        write_restart(rs_fn, k, k * dt, pot + ext, kin, crd, vel)

    # events: frames before step k (after k steps), restarts after step k (after k + 1 steps)
    marks = set()
    for iv in intvls:
        if iv > 0:
            marks.update(range(iv, n_steps + 1, iv))
    if rs_intvl > 0:
        marks.update(k + 1 for k in range(0, n_steps, rs_intvl))
    done = 0
    for m in sorted(x for x in marks if x <= n_steps) + [n_steps]:
        if m > done:
            t0 = time.perf_counter()
            system.run(m - done)
            t_step[0] = (time.perf_counter() - t0) / (m - done)
            done = m
        if m < n_steps:
            frame(m)
            if rs_intvl > 0 and (m - 1) % rs_intvl == 0:
                restart(m - 1)
    frame(n_steps)
    restart(n_steps)
    system.run(1)
    if langevin:
        restart(n_steps)
