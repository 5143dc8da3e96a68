"""Throughput of the MD force/integration path on one MI355X.

Workload (default): 33,000 rigid 3-site waters + one 200-bead flexible chain
(99,200 atoms, ~99.7 A cubic box), LJ + zero-dipole (alpha 0) with a 12 A
cutoff and a 13 A pair-list radius rebuilt every 10 steps, position Langevin
at 300 K (gamma 1/ps), dt 2 fs, SETTLE on every water.  One step = grid and
cell-pair rebuild (every 10th step) + pair kernel + excluded-pair/bond terms
+ fixed-point reduction + integration + constraints, all on the device,
captured as hipGraphs.  Inputs are resident in HBM before the timed region.

Printed JSON line: pair interactions/s (exact in-cutoff unmasked pair count
per step, counted by the pair kernel, times steps/s), ns/day beside it, the
pair-kernel roofline (algorithmic 41 FLOP per pair at alpha 0 over the
pair-kernel time measured with HIP events on the engine's stream) and the
CPU baseline (the oracle's single-thread cell-list force evaluation of the
same system).  Beside it ("alpha_large", skipped with --no-large): the pair
kernel at ZD alpha 0.1 on 400,000 waters (1.2 M atoms, rc 10 A, list 11 A);
and ("energies_at_print_interval", skipped with --no-sparse) the same step
with energies only every print-interval-log steps of the .cfg (1,000).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector, datasheet
FLOP_PER_PAIR = {0: 41, 1: 49}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_worker(ls, cfg, budget_s):
    """Oracle force evaluations (grid build + cell-pair list + FP32 LJ/ZD pair
    loop + excluded pairs) of one system until budget_s has elapsed."""
    import numpy as np

    import oracle
    zd = oracle.zd_setup(cfg["cutoff"], cfg["ele-alpha"])
    o = oracle.Oracle(ls.n_atoms, ls.box, cfg["nsgrid-cutoff"])
    tab = ls.nb15off_table()
    lj6, lj12 = ls.lj_tables_f32()
    q = ls.charge.astype(np.float32)
    ex = ls.excess_pairs()
    L, _, _ = oracle.pbc_from_box(ls.box)
    w_atom = np.zeros((ls.n_atoms, 3))
    n_eval, pairs = 0, 0
    t0 = time.perf_counter()
    while True:
        o.build(ls.crd)
        cps, _ = o.enumerate(tab, gap_bug=True)
        w, en, npair = o.pairs_fp32(zd, cps, ls.atom_type, lj6, lj12, q)
        oracle.excess(zd, L, ex, ls.crd, q, w_atom)
        n_eval += 1
        pairs += npair
        if time.perf_counter() - t0 >= budget_s:
            break
    return pairs, n_eval, time.perf_counter() - t0


def cpu_baseline(ls, cfg, budget_s=10.0, max_threads=16):
    """The oracle's cell-list force evaluation of the same system on the
    host: one thread for budget_s, then one independent evaluation loop per
    thread on every core this process may use (at most max_threads; the C
    oracle releases the GIL) for budget_s.  Returns the dict of both."""
    from concurrent.futures import ThreadPoolExecutor
    pairs, n_eval, dt = _cpu_worker(ls, cfg, budget_s)
    single = pairs / dt
    n_thr = max(1, min(max_threads, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(n_thr) as ex:
        res = list(ex.map(lambda _: _cpu_worker(ls, cfg, budget_s), range(n_thr)))
    multi = sum(p / t for p, _, t in res)
    n_multi = sum(n for _, n, _ in res)
    return {"value": multi, "unit": "pair interactions/s", "cores": n_thr, "kind": "port",
            "sample": "oracle single-point force evaluations (grid + cell-pair list + FP32 LJ/ZD pairs + excluded "
                      "pairs) of the same 99,200-atom system: %d threads x %.0f s (%d evaluations in total)"
                      % (n_thr, budget_s, n_multi),
            "single_thread": {"value": single, "evaluations": n_eval, "seconds": dt},
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0))}


# The walker and serving workloads build their cell-pair lists with the z-bin
# scan over every bin ("gap-bug": 0).  The reference's scan stops at the first
# uniform-z bin that no cell overlaps (MiniCell.cpp:1063-1077) and so drops
# interacting cell pairs wherever such a gap opens; over hundreds of steps
# atoms of a dropped pair drift into each other and the run ends in a NaN.
WALKER_CFG = {"cutoff": 12.0, "nsgrid-cutoff": 13.0, "nsgrid-update-intvl": 10, "time-step": 2.0, "gap-bug": 0,
              "integrator": "langevin", "constraint": "shake-settle", "const-tolerance": 1e-8,
              "langevin-gamma": 1.0, "temperature": 300.0, "ele-alpha": 0.0, "extended-ensemble": "vcmd",
              "enhance-sigma": 0.5, "enhance-recovery-coef": 1.0}


def _settle_waters(ls):
    import numpy as np
    ls.settle_ids, ls.settle_d = ls.shake_trio_ids, ls.shake_trio_d
    ls.shake_trio_ids = np.zeros((0, 3), np.int32)
    ls.shake_trio_d = np.zeros((0, 3))
    return ls


def equilibrate(systems_cfgs, steps=300, gamma=20.0):
    """Untimed relaxation of freshly generated systems: the seeded lattice
    waters start far from equilibrium and would heat to ~1,200 K.  Runs
    position Langevin with a strong friction (gamma ps^-1) on all systems
    in one batch and returns launch sets carrying the relaxed coordinates
    and velocities."""
    import copy
    from omegagene_amd.engine import Batch, System
    systems = [System(ls, dict(c, **{"langevin-gamma": gamma, "extended-ensemble": "none"}))
               for ls, c in systems_cfgs]
    b = Batch()
    for x in systems:
        b.add(x)
    b.run(steps)
    out = []
    for (ls, _), x in zip(systems_cfgs, systems):
        crd, _, vj = x.state()
        ls2 = copy.copy(ls)
        ls2.crd, ls2.vel = crd, vj
        out.append(ls2)
    b.close()
    for x in systems:
        x.close()
    return out


def bench_large(steps=30, warmup=10, split=False):
    """400,000 waters (1,200,000 atoms, 228.7 A box), rc 10 A, list 11 A, ZD
    alpha 0.1: the pair kernel at alpha > 0 timed with HIP events over MD
    steps (Langevin 300 K at gamma 20/ps straight from the seeded lattice,
    SETTLE, rebuild every 10 steps)."""
    from omegagene_amd import gen
    from omegagene_amd.engine import System, lib
    ls, cfg = gen.config2()
    ls = _settle_waters(ls)
    cfg = dict(cfg, **{"integrator": "langevin", "langevin-gamma": 20.0, "log-cap": steps,
                       "split-step": int(split)})
    s = System(ls, cfg)
    s.run(warmup)
    s.set_pair_timing(True)
    lib().og_thread_sync()
    t0 = time.perf_counter()
    s.run(steps)
    lib().og_thread_sync()
    wall = time.perf_counter() - t0
    pair_ms, n_launch = s.pair_time()
    elog = s.energy_log(max_rows=steps)
    s.close()
    pairs = float(elog[:, 8].mean())
    t_launch = pair_ms / max(n_launch, 1) * 1e-3
    tflops = FLOP_PER_PAIR[1] * pairs / t_launch / 1e12
    log("1.2 M atoms, alpha 0.1: %d steps in %.3f s, pair kernel %.3f ms" % (steps, wall, t_launch * 1e3))
    return {"workload": "400,000 waters, 1,200,000 atoms, 228.7 A box, rc 10 A, list 11 A, ZD alpha 0.1, "
                        "Langevin 300 K (gamma 20/ps from the lattice), SETTLE, rebuild every 10 steps",
            "pairs_per_step": pairs, "ms_per_step": wall / steps * 1e3,
            "pair_interactions_per_s": pairs * steps / wall, "pair_kernel_ms": t_launch * 1e3,
            "pair_kernel_tflops": tflops, "flop_per_pair": FLOP_PER_PAIR[1],
            "frac_of_fp32_peak": tflops / PEAK_FP32_TFLOPS, "timed_steps": steps}


def bench_nve(steps=1000, relax=300, gap_bug=1):
    """NVE energy drift per configuration: each system is relaxed (position
    Langevin, gamma 20/ps, SETTLE or SHAKE as configured) and then run
    `steps` leapfrog steps at dt 2 fs with energies logged every step.  Drift
    is the slope of a least-squares line through the total energy
    (potential + kinetic), in kcal/mol per ns and per atom per ns; the
    fluctuation is the RMS residual about that line."""
    import numpy as np
    from omegagene_amd import gen
    from omegagene_amd.engine import EngineError, System, potential
    out = []
    cases = [("6,000 atoms (2,000 waters), rc 12 A, alpha 0, SHAKE", gen.config0(), False),
             ("99,200 atoms (33,000 waters + 200-bead chain), rc 12 A, alpha 0, SETTLE", gen.config1(), True),
             ("1,200,000 atoms (400,000 waters), rc 10 A, alpha 0.1, SETTLE", gen.config2(), True)]
    for name, (ls, cfg), settle in cases:
        if settle:
            ls = _settle_waters(ls)
        cfg = dict(cfg, **{"constraint": "shake-settle" if settle else "shake", "integrator": "langevin",
                           "temperature": 300.0, "gap-bug": gap_bug})
        rule = "reference z-bin scan" if gap_bug else "full z-bin scan"
        try:
            ls = equilibrate([(ls, cfg)], steps=relax)[0]
            nve = dict(cfg, **{"integrator": "leapfrog-presto", "log-cap": steps})
            s = System(ls, nve)
            try:
                s.run(steps)
                el = s.energy_log(max_rows=steps)
            finally:
                s.close()
        except EngineError as e:
            out.append({"workload": name, "n_atoms": int(ls.n_atoms), "list_rule": rule, "error": str(e)})
            log("nve %s (%s): %s" % (name, rule, e))
            continue
        tot = potential(el) + el[:, 5] + el[:, 7]
        t_ns = np.arange(len(tot)) * cfg["time-step"] * 1e-6
        slope, icpt = np.polyfit(t_ns, tot, 1)
        resid = tot - (slope * t_ns + icpt)
        out.append({"workload": name, "n_atoms": int(ls.n_atoms), "steps": steps, "dt_fs": cfg["time-step"],
                    "list_rule": rule,
                    "drift_kcal_per_mol_per_ns": float(slope),
                    "drift_kcal_per_mol_per_ns_per_atom": float(slope / ls.n_atoms),
                    "fluctuation_kcal_per_mol": float(resid.std()), "mean_kinetic_kcal_per_mol": float(el[:, 5].mean()),
                    "total_start": float(tot[0]), "total_end": float(tot[-1])})
        r = out[-1]
        log("nve %s (%s): drift %.3f kcal/mol/ns (%.2e per atom), fluctuation %.3f" %
            (name, r["list_rule"], r["drift_kcal_per_mol_per_ns"], r["drift_kcal_per_mol_per_ns_per_atom"],
             r["fluctuation_kcal_per_mol"]))
    return out


def call_point_walker_rate(walkers, cfgs, steps=40):
    """The walkers one at a time through the reference's own call points
    (include/og_celeste.h, the plain drop-in path, one system per process as
    in the reference): per step the grid-order coordinates go up
    (og_memcpy_htod_crd + og_set_crd), the pair kernel runs (og_pairwise_ljzd
    + og_pair_sync) and work + energies come back (og_memcpy_dtoh_work);
    every nsgrid-update-intvl steps the grid is sent again
    (og_set_cell_constant ... og_enumerate_cell_pairs).  The grid is the
    engine's own (og_get_grid).  The host side of the reference's step
    (MiniCell rebuild, integration, constraints) is not run, so this rate is
    an upper bound for that path."""
    import numpy as np
    from omegagene_amd.celeste import SubBoxDriver, ZeroDipole, pbc_from_box
    from omegagene_amd.engine import System
    n_done, t_total = 0, 0.0
    for w, c in zip(walkers, cfgs):
        s = System(w, c)
        s.nsgrid_update()
        ids, gcrd, head = s.grid()
        geo = s.grid_geometry()
        n_cells = s.grid_info()["n_cells"]
        s.close()
        L, lb = pbc_from_box(w.box)
        lj6, lj12 = w.lj_tables_f32()
        drv = SubBoxDriver(w.n_atoms, L, lb, c["cutoff"], c["nsgrid-cutoff"], ZeroDipole(c["cutoff"]), lj6, lj12,
                           w.nb15off_table(), w.charge.astype(np.float32), w.atom_type, geo,
                           max_n_atom_array=geo["max_n_cells"] * 8)
        grid = {"n_cells": n_cells, "atomids": ids, "crd": gcrd.astype(np.float32), "idx_xy_head_cell": head,
                "column_of_slot": np.zeros(len(ids), np.int32)}
        intvl = int(c["nsgrid-update-intvl"])
        try:
            drv.update_grid(grid)
            drv.pair_forces()
            t0 = time.perf_counter()
            for k in range(steps):
                if k % intvl == 0:
                    drv.update_grid(grid)
                else:
                    drv.set_coordinates(grid["crd"])
                drv.pair_forces()
            t_total += time.perf_counter() - t0
            n_done += steps
        finally:
            drv.close()
    return n_done / t_total


def bench_walkers(n_walkers=64, steps=500, seq_steps=100, warmup=50):
    """64 independent water + 200-bead-chain walkers (3,000-30,000 atoms,
    log-uniform), Langevin 300 K, SETTLE, VcMD bias on the chain end-to-end
    distance (8 windows).  Batched: one og_batch, one 500-step hipGraph per
    launch, MiniCell column width from the cell-edge autotuner.  Baselines:
    the same walkers one at a time through the single-system path
    (og_md_run), seq_steps each, and through the reference's call points."""
    import numpy as np
    from omegagene_amd import gen
    from omegagene_amd.engine import Batch, System, lib
    sizes = gen.walker_sizes(n_walkers)
    walkers = [_settle_waters(gen.walker_system(int(n), seed=5000 + i)) for i, n in enumerate(sizes)]
    cfgs = [dict(WALKER_CFG, **{"random-seed": 100 + i, "log-cap": steps + warmup}) for i in range(n_walkers)]
    walkers = equilibrate(list(zip(walkers, cfgs)))
    n_atoms = sum(w.n_atoms for w in walkers)
    # the MiniCell column width, tuned on the median walker (the full z-bin scan lets it vary)
    from omegagene_amd.engine import autotune_cell_edge
    mid = int(np.argsort([w.n_atoms for w in walkers])[len(walkers) // 2])
    edge, edge_ms = autotune_cell_edge(walkers[mid], cfgs[mid], trials=3)
    cfgs = [dict(c, **{"cell-edge-scale": edge}) for c in cfgs]
    log("walkers: cell edge x%.2f of the reference width (%.1f us pair kernel on the %d-atom walker)" %
        (edge, edge_ms * 1e3, walkers[mid].n_atoms))
    log("walkers: %d systems, %d atoms in total (%d-%d)" % (n_walkers, n_atoms, min(sizes), max(sizes)))
    # one at a time (the plain drop-in path)
    seq_time, seq_steps_done = 0.0, 0
    for w, c in zip(walkers, cfgs):
        s = System(w, c)
        s.run(10)
        lib().og_thread_sync()
        t0 = time.perf_counter()
        s.run(seq_steps)
        lib().og_thread_sync()
        seq_time += time.perf_counter() - t0
        seq_steps_done += seq_steps
        s.close()
    seq_rate = seq_steps_done / seq_time
    log("walkers one at a time: %.1f walker-steps/s" % seq_rate)
    cp_rate = call_point_walker_rate(walkers, cfgs)
    log("walkers one at a time through the call points: %.1f walker-steps/s (device part only)" % cp_rate)
    # batched
    systems = [System(w, c) for w, c in zip(walkers, cfgs)]
    b = Batch()
    for s in systems:
        b.add(s)
    lib().og_batch_set_graph_steps(b.h, steps)
    b.run(warmup)
    tune = b.autotune(5)
    log("walkers: autotuned pair kernel %d waves per block, block order %s" % (tune["waves_per_block"],
                                                                              tune["order"]))
    lib().og_thread_sync()
    t0 = time.perf_counter()
    b.run(steps)
    lib().og_thread_sync()
    wall = time.perf_counter() - t0
    batch_rate = n_walkers * steps / wall
    log("walkers batched: %.1f walker-steps/s (%.2f s for %d steps)" % (batch_rate, wall, steps))
    b.close()
    for s in systems:
        s.close()
    return {"workload": "64 walkers, water + 200-bead chain, 3,000-30,000 atoms log-uniform (%d atoms in total), "
                        "Langevin 300 K, SETTLE, VcMD 8 windows, rc 12 A" % n_atoms,
            "batched_walker_steps_per_s": batch_rate, "one_at_a_time_walker_steps_per_s": seq_rate,
            "speedup": batch_rate / seq_rate,
            "one_at_a_time_call_points_walker_steps_per_s": cp_rate,
            "speedup_vs_call_points": batch_rate / cp_rate,
            "baselines": "speedup: against the same walkers one at a time on the device-resident single-system "
                         "path (og_md_run); speedup_vs_call_points: against the device part of the reference's "
                         "call-point path (og_celeste.h) one walker at a time, host work not counted",
            "atom_steps_per_s": n_atoms * steps / wall,
            "graph_steps": steps, "timed_steps": steps, "pair_shape": dict(tune, cell_edge_scale=edge)}


def bench_serving(n_requests=48, rate=200.0, budget_gb=200.0, chunk=20, seed=7):
    """Poisson stream (mean 200 requests/s) of segments: each a new walker of
    2,000-50,000 atoms (log-uniform, water + 200-bead chain), 200-2,000 steps,
    280-360 K, admitted against a 200 GB HBM budget.  Latency = completion
    time - arrival time."""
    import numpy as np
    from omegagene_amd import gen
    from omegagene_amd.engine import Server
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    sizes = np.exp(rng.uniform(np.log(2000), np.log(50000), n_requests)).astype(int)
    nsteps = rng.integers(200, 2001, n_requests)
    temps = rng.uniform(280.0, 360.0, n_requests)
    systems = [_settle_waters(gen.walker_system(int(n), seed=7000 + i, vcmd=False)) for i, n in enumerate(sizes)]
    cfg = {k: v for k, v in WALKER_CFG.items() if k not in ("extended-ensemble", "enhance-sigma", "enhance-recovery-coef")}
    srv = Server(int(budget_gb * (1 << 30)), chunk_steps=chunk)
    t_start = srv.stats()["t"]
    reqs = {}
    nxt = 0
    while True:
        now = srv.stats()["t"] - t_start
        while nxt < n_requests and arrivals[nxt] <= now:
            c = dict(cfg, **{"temperature": float(temps[nxt]), "random-seed": 900 + nxt})
            rid, _ = srv.submit_new(systems[nxt], c, int(nsteps[nxt]))
            reqs[rid] = nxt
            nxt += 1
        active, queued, _ = srv.pump()
        if nxt == n_requests and active == 0 and queued == 0:
            break
        if active == 0 and queued == 0:
            time.sleep(max(0.0, min(0.001, arrivals[nxt] - now)))
    lat = []
    for rid, k in reqs.items():
        _, _, _, t = srv.result(rid)
        lat.append(t[2] - t_start - arrivals[k])
    st = srv.stats()
    wall = st["t"] - t_start
    srv.close()
    lat = np.array(lat)
    atom_steps = float(np.sum(sizes * nsteps))
    return {"workload": "%d requests, Poisson %.0f req/s, 2,000-50,000 atoms, 200-2,000 steps, 280-360 K, "
                        "200 GB budget, %d-step chunks" % (n_requests, rate, chunk),
            "p50_latency_ms": float(np.percentile(lat, 50) * 1e3), "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
            "aggregate_steps_per_s": float(nsteps.sum() / wall), "atom_steps_per_s": atom_steps / wall,
            "wall_s": wall}


def copy_calibration(mib: int):
    """A plain device copy of known size (mib MiB read + mib MiB written), run
    under the same rocprofv3 counters as the engine to scale FETCH_SIZE /
    WRITE_SIZE into bytes."""
    import torch
    x = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "bytes_read_per_copy": mib << 20, "bytes_written_per_copy": mib << 20}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-walkers", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-sparse", action="store_true",
                    help="skip the run with energies only at the .cfg print interval")
    ap.add_argument("--large-only", action="store_true",
                    help="only the 1.2 M-atom alpha 0.1 measurement (its own JSON line; for profiling)")
    ap.add_argument("--nve-drift", action="store_true",
                    help="only the NVE energy drift of the three single-system configurations (own JSON line)")
    ap.add_argument("--split-step", action="store_true",
                    help="run the integration step as separate kernels (for profiling them)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.nve_drift:
        print(json.dumps({"nve_drift": bench_nve(gap_bug=0) + bench_nve(gap_bug=1)}))
        return
    if args.large_only:
        print(json.dumps({"alpha_large": bench_large(steps=args.steps, warmup=args.warmup,
                                                     split=args.split_step)}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs each system (or walker batch) on a single "
              "MI355X; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from omegagene_amd import gen
    from omegagene_amd.engine import System, lib

    ls, cfg = gen.config1()
    cfg = dict(cfg)
    # every water through SETTLE (rigid-body analytic constraint), the chain is flexible
    ls.settle_ids, ls.settle_d = ls.shake_trio_ids, ls.shake_trio_d
    import numpy as np
    ls.shake_trio_ids = np.zeros((0, 3), np.int32)
    ls.shake_trio_d = np.zeros((0, 3))
    cfg["log-cap"] = max(args.steps, 1)
    cfg["split-step"] = int(args.split_step)
    log("system: %d atoms, box %.3f A, cutoff %.1f / list %.1f, rebuild every %d steps" %
        (ls.n_atoms, ls.box[0], cfg["cutoff"], cfg["nsgrid-cutoff"], cfg["nsgrid-update-intvl"]))
    ls = equilibrate([(ls, cfg)])[0]
    log("equilibrated 300 steps at gamma 20/ps (untimed)")
    s = System(ls, cfg)
    alpha_mode = 1 if cfg["ele-alpha"] > 0 else 0
    s.run(args.warmup)
    log("warmup %d steps done" % args.warmup)
    tune = s.autotune(5)
    log("autotuned pair kernel: %d waves per block (%.3f ms)" % (tune["waves_per_block"], tune["pair_ms"]))

    # the timed region: the production path (hipGraphs of K steps, no per-launch events)
    lib().og_thread_sync()
    t0 = time.perf_counter()
    s.run(args.steps)
    lib().og_thread_sync()
    wall = time.perf_counter() - t0
    elog = s.energy_log(max_rows=args.steps)
    log("timed %d steps in %.3f s (graph path)" % (args.steps, wall))
    # the pair-kernel time from a separate pass of the same system: steps launched directly with HIP
    # events around every pair-kernel launch on the engine's stream
    n_pt = min(args.steps, 500)
    s.set_pair_timing(True)
    lib().og_thread_sync()
    t_pt = time.perf_counter()
    s.run(n_pt)
    lib().og_thread_sync()
    wall_pt = time.perf_counter() - t_pt
    pair_ms, n_launch = s.pair_time()
    s.set_pair_timing(False)
    log("pair-kernel timing pass: %d steps, %.4f ms per pair launch, %.4f ms/step without graphs" %
        (n_pt, pair_ms / max(n_launch, 1), wall_pt / n_pt * 1e3))

    npairs = elog[:, 8]
    pairs_per_step = float(npairs.mean())
    steps_per_s = args.steps / wall
    value = pairs_per_step * steps_per_s
    ns_day = steps_per_s * cfg["time-step"] * 86400.0 / 1e6
    t_launch = pair_ms / max(n_launch, 1) * 1e-3
    flops = FLOP_PER_PAIR[alpha_mode] * pairs_per_step
    achieved = flops / t_launch / 1e12 if t_launch > 0 else 0.0
    T = elog[-1, 5] * 2 * 4184.0 / (8.31451 * s.d_free)

    # the same workload with energies only at the .cfg's print interval (the trajectory is the same; the pair
    # kernel then skips the energy terms on the other steps): a second system continued from the final state
    sparse = None
    if not args.no_sparse:
        crd, _, vj = s.state()
        import copy
        ls2 = copy.copy(ls)
        ls2.crd, ls2.vel = crd, vj
        stride = int(cfg.get("print-interval-log", 1000))
        s2 = System(ls2, dict(cfg, **{"log-stride": stride, "log-cap": max(args.steps // stride, 1)}))
        s2.run(args.warmup)
        s2.set_pair_shape(tune["waves_per_block"], tune["order"])
        lib().og_thread_sync()
        t1 = time.perf_counter()
        s2.run(args.steps)
        lib().og_thread_sync()
        wall2 = time.perf_counter() - t1
        s2.close()
        sparse = {"energies_every_steps": stride, "ms_per_step": wall2 / args.steps * 1e3,
                  "pair_interactions_per_s": pairs_per_step * args.steps / wall2,
                  "ns_per_day": args.steps / wall2 * cfg["time-step"] * 86400.0 / 1e6,
                  "note": "pairs per step from the energies-every-step run (same trajectory rule)"}
        log("energies every %d steps: %.4f ms/step" % (stride, sparse["ms_per_step"]))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(ls, cfg)
        log("cpu baseline: %.3e pair interactions/s on %d threads (%.3e on one), %s" %
            (cpu["value"], cpu["cores"], cpu["single_thread"]["value"], cpu["cpu_model"]))

    # HBM bytes per pair-kernel launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE passes
    # (a counter run cannot be nested in this process)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_pair_water_chain_99k_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath))["traffic_bytes_per_launch"]
    # FP32 work the pair kernel issues per launch, from the committed SQ_INSTS_VALU_*_F32 counter pass
    issued = None
    vpath = os.path.join(ROOT, "profiles", "k_pair_water_chain_99k_valu.json")
    if os.path.exists(vpath) and t_launch > 0:
        issued = json.load(open(vpath))["issued_fp32_flop_per_launch"] / t_launch / 1e12
    walkers = None if args.no_walkers else bench_walkers()
    serving = None if args.no_serving else bench_serving()
    large = None if args.no_large else bench_large()

    out = {
        "metric": "pair interactions/s (99,200 atoms, 12 A, ZD alpha 0, Langevin, dt 2 fs)",
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pairs, int64 fixed-point forces, fp64 integration",
        "data": "synthetic (seeded water lattice + 200-bead chain)",
        "config": {"workload": "33,000 waters + 200-bead chain, 99,200 atoms, 99.7 A box, rc 12 A, list 13 A, "
                               "rebuild every 10 steps, Langevin 300 K, SETTLE; water H carries LJ (sigma 0.4 A, "
                               "eps 0.046) beside the charged chain, so a bead cannot collapse onto an H",
                   "n_atoms": ls.n_atoms, "pairs_per_step": pairs_per_step},
        "ns_per_day": ns_day,
        "temperature_K": T,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel": "k_pair<%d, %d>" % (alpha_mode, tune["waves_per_block"]), "kernel_ms": t_launch * 1e3,
                     "flop_per_pair": FLOP_PER_PAIR[alpha_mode],
                     "issued_tflops_counters": issued,
                     "timing": "HIP events around each pair-kernel launch over %d further steps launched without "
                               "graphs (%d launches); the headline steps run as hipGraphs" % (n_pt, n_launch)},
        "cpu_baseline": cpu,
        "walkers": walkers,
        "serving": serving,
        "alpha_large": large,
        "energies_at_print_interval": sparse,
    }
    s.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
