"""Seeded synthetic systems for the five benchmark workloads, written as
``.cls`` launch sets plus ``.cfg`` files with the reference's keys.

Water: rigid 3-site (O-H 0.9572 A, H-O-H 104.52 deg), LJ on O
(sigma 3.1507 A, epsilon 0.1521 kcal/mol; in systems with the charged chain
also a small hydrogen LJ, sigma 0.4 A, epsilon 0.046, so that a bead cannot
collapse onto a hydrogen), charges O -0.834 / H +0.417, on a
jittered cubic lattice (+-0.1 spacing) with uniformly random orientations.
Chain: 200 beads on a seeded self-avoiding walk with 1.53 A bonds, harmonic
E = k (r - r0)^2 with k = 300 kcal/mol/A^2 (the reference convention has no
1/2, ForceField.cpp:108-136), per-bead LJ sigma ~ U(3, 4) A and
epsilon ~ U(0.05, 0.2) kcal/mol with one LJ type per bead and
Lorentz-Berthelot cross terms against water O (a builder's choice), charges
+-0.5 in equal numbers (net zero), bead mass 14.027 amu (a CH2 unit),
exclusions 1-2, 1-3, 1-4.  Boxes are cubic at 1.0 g/cm^3.  LJ tables use the
reference convention E = A12/r^12 - A6/r^6 with A6 = 4 eps sigma^6 and
A12 = 4 eps sigma^12.  Velocities are Maxwell-Boltzmann at T, with the
centre-of-mass motion removed and the components along each rigid water's
constraints projected out, then rescaled to T over the constrained degrees
of freedom (DynamicsMode.cpp:61-67 convention).
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

from .cls import LaunchSet, VcmdDef

D_OH = 0.9572
THETA_HOH = 104.52
D_HH = 2.0 * D_OH * math.sin(math.radians(THETA_HOH) / 2.0)
SIG_O, EPS_O = 3.1507, 0.1521
# water hydrogen LJ of the CHARMM TIP3P variant (Rmin/2 0.2245 A): without it a
# charged solute site can collapse onto a hydrogen (no repulsion against the
# Coulomb attraction)
SIG_H, EPS_H = 0.4, 0.046
Q_O, Q_H = -0.834, 0.417
M_O, M_H = 15.9994, 1.008
M_BEAD = 14.027
BOND_R0, BOND_K = 1.53, 300.0
AMU_A3_TO_GCM3 = 1.66053907      # 1 amu/A^3 in g/cm^3
GAS_CONST = 8.31451              # CelesteObject.cpp:38
KINETIC_COEFF = (1e7 / (4.184 * 1e3)) * 0.5


def box_for_density(total_mass_amu: float, density: float = 1.0) -> float:
    return (total_mass_amu * AMU_A3_TO_GCM3 / density) ** (1.0 / 3.0)


def _random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _mimg(d, L):
    return d - L * np.round(d / L)


def _chain_walk(rng, n_beads, L, min_sep=3.6, max_tries=400, bond=BOND_R0):
    """Self-avoiding walk (minimum image) with the given bond (1.53 A) and bend
    <= 70 deg; beads separated by 4+ bonds stay >= min_sep apart."""
    pos = np.zeros((n_beads, 3))
    pos[0] = rng.uniform(0, L, 3)
    prev_dir = None
    i = 1
    fails = 0
    while i < n_beads:
        ok = False
        for _ in range(max_tries):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            if prev_dir is not None and np.dot(u, prev_dir) < math.cos(math.radians(70.0)):
                continue
            p = pos[i - 1] + bond * u
            if i >= 4:
                d = _mimg(pos[: i - 3] - p, L)
                if np.min(np.einsum("ij,ij->i", d, d)) < min_sep * min_sep:
                    continue
            ok = True
            break
        if ok:
            pos[i] = p
            prev_dir = u
            i += 1
        else:
            fails += 1
            back = min(i - 1, 5)
            i -= back
            prev_dir = None if i < 2 else (pos[i - 1] - pos[i - 2]) / BOND_R0
            if fails > 2000:
                raise RuntimeError("chain walk failed")
    return pos


def build_system(n_waters: int, seed: int, L: Optional[float] = None, n_beads: int = 0, T: float = 300.0,
                 with_vcmd: bool = False, h_lj: bool = True) -> LaunchSet:
    """Water (+ optional chain) system at 1.0 g/cm^3 unless L is given.
    h_lj=False leaves the water hydrogens without LJ (LJ on O only, the
    plain-water workloads); with a charged chain the hydrogens keep the small
    CHARMM TIP3P LJ so that a bead cannot collapse onto one."""
    rng = np.random.default_rng(seed)
    total_mass = n_waters * (M_O + 2 * M_H) + n_beads * M_BEAD
    if L is None:
        L = box_for_density(total_mass)
    # chain first, then lattice sites away from it
    beads = _chain_walk(rng, n_beads, L) if n_beads else np.zeros((0, 3))
    m = int(math.ceil(n_waters ** (1.0 / 3.0)))
    while True:
        a = L / m
        g = (np.arange(m) + 0.5) * a
        sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
        if n_beads:
            keep = np.ones(len(sites), bool)
            for s in range(0, len(sites), 4096):
                d = _mimg(sites[s:s + 4096, None, :] - beads[None], L)
                keep[s:s + 4096] = np.min(np.einsum("ijk,ijk->ij", d, d), 1) >= 3.3 ** 2
            sites = sites[keep]
        if len(sites) >= n_waters:
            break
        m += 1
    sel = rng.choice(len(sites), n_waters, replace=False)
    sel.sort()
    O = sites[sel] + rng.uniform(-0.1, 0.1, (n_waters, 3)) * a
    half = math.radians(THETA_HOH) / 2.0
    body = np.array([[0.0, 0.0, 0.0],
                     [D_OH * math.sin(half), D_OH * math.cos(half), 0.0],
                     [-D_OH * math.sin(half), D_OH * math.cos(half), 0.0]])
    R = _random_rotations(rng, n_waters)
    wat = O[:, None, :] + np.einsum("nij,kj->nki", R, body)
    n = 3 * n_waters + n_beads
    crd = np.concatenate([wat.reshape(-1, 3), beads], 0)
    crd = np.mod(crd, L)
    crd[crd >= L] -= L  # guard against mod rounding onto L

    # topology
    n_types = 2 + n_beads
    sig = np.zeros(n_types)
    eps = np.zeros(n_types)
    sig[0], eps[0] = SIG_O, EPS_O
    if h_lj:
        sig[1], eps[1] = SIG_H, EPS_H
    if n_beads:
        sig[2:] = rng.uniform(3.0, 4.0, n_beads)
        eps[2:] = rng.uniform(0.05, 0.2, n_beads)
    s_ij = 0.5 * (sig[:, None] + sig[None, :])
    e_ij = np.sqrt(eps[:, None] * eps[None, :])
    lj6 = 4.0 * e_ij * s_ij ** 6
    lj12 = 4.0 * e_ij * s_ij ** 12
    atom_type = np.empty(n, np.int32)
    atom_type[0:3 * n_waters:3] = 0
    atom_type[1:3 * n_waters:3] = 1
    atom_type[2:3 * n_waters:3] = 1
    atom_type[3 * n_waters:] = 2 + np.arange(n_beads)
    charge = np.empty(n)
    charge[0:3 * n_waters:3] = Q_O
    charge[1:3 * n_waters:3] = Q_H
    charge[2:3 * n_waters:3] = Q_H
    if n_beads:
        q = np.array([0.5] * (n_beads // 2) + [-0.5] * (n_beads - n_beads // 2))
        rng.shuffle(q)
        q -= q.mean()
        charge[3 * n_waters:] = q
    mass = np.empty(n)
    mass[0:3 * n_waters:3] = M_O
    mass[1:3 * n_waters:3] = M_H
    mass[2:3 * n_waters:3] = M_H
    mass[3 * n_waters:] = M_BEAD

    w = np.arange(n_waters, dtype=np.int32) * 3
    excl = [np.stack([w, w + 1], 1), np.stack([w, w + 2], 1), np.stack([w + 1, w + 2], 1)]
    b0 = 3 * n_waters
    for k in (1, 2, 3):
        if n_beads > k:
            i = np.arange(n_beads - k, dtype=np.int32) + b0
            excl.append(np.stack([i, i + k], 1))
    nb15 = np.concatenate(excl, 0)
    # reference order: per atom, partners grouped; sort by (low, high) for a stable input order
    nb15 = nb15[np.lexsort((nb15[:, 1], nb15[:, 0]))].astype(np.int32)

    box = np.zeros(12)
    box[0] = box[4] = box[8] = L
    box[9:12] = 0.5 * L

    ls = LaunchSet(box=box, crd=crd, vel=np.zeros_like(crd), charge=charge, mass=mass, atom_type=atom_type,
                   n_lj_types=n_types, lj6=lj6, lj12=lj12, nb15off=nb15)
    if n_beads:
        bi = np.arange(n_beads - 1, dtype=np.int32) + b0
        ls.bond_ids = np.stack([bi, bi + 1], 1)
        ls.bond_k = np.full(n_beads - 1, BOND_K)
        ls.bond_r0 = np.full(n_beads - 1, BOND_R0)
    ls.shake_trio_ids = np.stack([w, w + 1, w + 2], 1).astype(np.int32)
    ls.shake_trio_d = np.tile([D_OH, D_HH, D_OH], (n_waters, 1))
    ls.groups = [("solvent", np.arange(3 * n_waters, dtype=np.int32))]
    if n_beads:
        ls.groups.insert(0, ("solute", np.arange(b0, n, dtype=np.int32)))
        if with_vcmd:
            ls.groups.append(("end_a", np.arange(b0, b0 + 5, dtype=np.int32)))
            ls.groups.append(("end_b", np.arange(n - 5, n, dtype=np.int32)))
    ls.vel = maxwell_velocities(ls, T, rng)
    if with_vcmd and n_beads:
        ls.vcmd = vcmd_windows(ls, n_windows=8, seed=seed)
    return ls


# coarse-grained chains for the HPS + Debye-Hueckel pair form (one bead per
# residue, as the reference's cg_q8 sample): bond 3.8 A with k = 10
# kcal/mol/A^2, bead mass 110 amu, HPS epsilon 0.2 kcal/mol
CG_BOND_R0, CG_BOND_K, CG_MASS, HPS_EPS = 3.8, 10.0, 110.0, 0.2


def build_cg_system(n_chains: int, n_beads: int, seed: int, L: float, n_types: int = 20,
                    T: float = 300.0) -> LaunchSet:
    """Coarse-grained chains in a cubic box of edge L: per-type sigma
    ~ U(4.5, 7) A and hydrophobicity ~ U(0, 1); types 0-2 carry +1, types 3-5
    carry -1, the rest are neutral.  LJ A6 = 4 eps s^6, A12 = 4 eps s^12 at
    eps = HPS_EPS with s the mean sigma of the pair; the HPS section holds
    cutoff 2^(1/6) s and lambda the mean hydrophobicity.  Exclusions: bonded
    neighbours (1-2)."""
    rng = np.random.default_rng(seed)
    pos = []
    for _ in range(n_chains):
        for _ in range(200):
            c = _chain_walk(rng, n_beads, L, min_sep=4.5, bond=CG_BOND_R0)
            if not pos:
                break
            prev = np.concatenate(pos, 0)
            d = _mimg(c[:, None, :] - prev[None], L)
            if np.min(np.einsum("ijk,ijk->ij", d, d)) >= 4.5 ** 2:
                break
        else:
            raise RuntimeError("chain placement failed")
        pos.append(c)
    crd = np.mod(np.concatenate(pos, 0), L)
    crd[crd >= L] -= L
    n = len(crd)
    sig_t = rng.uniform(4.5, 7.0, n_types)
    lam_t = rng.uniform(0.0, 1.0, n_types)
    q_t = np.zeros(n_types)
    q_t[0:3], q_t[3:6] = 1.0, -1.0
    atom_type = rng.integers(0, n_types, n).astype(np.int32)
    s_ij = 0.5 * (sig_t[:, None] + sig_t[None, :])
    lj6 = 4.0 * HPS_EPS * s_ij ** 6
    lj12 = 4.0 * HPS_EPS * s_ij ** 12
    box = np.zeros(12)
    box[0] = box[4] = box[8] = L
    box[9:12] = 0.5 * L
    ls = LaunchSet(box=box, crd=crd, vel=np.zeros_like(crd), charge=q_t[atom_type].copy(), mass=np.full(n, CG_MASS),
                   atom_type=atom_type, n_lj_types=n_types, lj6=lj6, lj12=lj12,
                   hps_cutoff=2.0 ** (1.0 / 6.0) * s_ij, hps_lambda=0.5 * (lam_t[:, None] + lam_t[None, :]))
    bi = np.concatenate([np.arange(n_beads - 1) + k * n_beads for k in range(n_chains)]).astype(np.int32)
    ls.bond_ids = np.stack([bi, bi + 1], 1)
    ls.bond_k = np.full(len(bi), CG_BOND_K)
    ls.bond_r0 = np.full(len(bi), CG_BOND_R0)
    ls.nb15off = ls.bond_ids.copy()
    ls.groups = [("all", np.arange(n, dtype=np.int32))]
    ls.vel = maxwell_velocities(ls, T, rng)
    return ls


def _project_rigid_water(crd, vel, mass, ids, L):
    """Remove velocity components along the three constraints of each water."""
    a, b, c = ids[:, 0], ids[:, 1], ids[:, 2]
    pairs = [(a, b), (b, c), (c, a)]
    d = [_mimg(crd[i] - crd[j], L) for i, j in pairs]
    minv = 1.0 / mass
    A = np.zeros((len(ids), 3, 3))
    rhs = np.zeros((len(ids), 3))
    for k, (i, j) in enumerate(pairs):
        rhs[:, k] = -np.einsum("nd,nd->n", d[k], vel[i] - vel[j])
        for l, (p, q) in enumerate(pairs):
            # d_k . (dv_i - dv_j) where dv_x = sum_l lam_l d_l (+minv_x if x==p, -minv_x if x==q)
            coef = np.zeros(len(ids))
            coef += np.where(i == p, minv[i], 0.0) - np.where(i == q, minv[i], 0.0)
            coef -= np.where(j == p, minv[j], 0.0) - np.where(j == q, minv[j], 0.0)
            A[:, k, l] = coef * np.einsum("nd,nd->n", d[k], d[l])
    lam = np.linalg.solve(A, rhs[..., None])[..., 0]
    out = vel.copy()
    for l, (p, q) in enumerate(pairs):
        out[p] += (lam[:, l] * minv[p])[:, None] * d[l]
        out[q] -= (lam[:, l] * minv[q])[:, None] * d[l]
    return out


def maxwell_velocities(ls: LaunchSet, T: float, rng) -> np.ndarray:
    m = ls.mass
    sd = np.sqrt(GAS_CONST * T / (m * 1e-3)) * 1e-5   # m/s -> A/fs
    v = rng.normal(size=(ls.n_atoms, 3)) * sd[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    if len(ls.shake_trio_ids):
        v = _project_rigid_water(ls.crd, v, m, ls.shake_trio_ids, ls.box[0])
    v -= (m[:, None] * v).sum(0) / m.sum()
    ke = KINETIC_COEFF * np.sum(m[:, None] * v * v)
    t_now = ke * 2.0 * 4184.0 / (GAS_CONST * ls.d_free(constraint=True))
    v *= math.sqrt(T / t_now)
    return v


def end_to_end(ls: LaunchSet) -> float:
    ga = dict(ls.groups)["end_a"]
    gb = dict(ls.groups)["end_b"]
    ca = (ls.crd[ga] * ls.mass[ga, None]).sum(0) / ls.mass[ga].sum()
    cb = (ls.crd[gb] * ls.mass[gb, None]).sum(0) / ls.mass[gb].sum()
    return float(np.linalg.norm(_mimg(ca - cb, ls.box[0])))


def vcmd_windows(ls: LaunchSet, n_windows: int = 8, seed: int = 0) -> VcmdDef:
    """1-D VcMD over the chain end-to-end distance: 8 overlapping windows
    between 4 A and half the box edge, starting in the window holding the
    initial distance (Read.cpp:705-779 layout)."""
    L = ls.box[0]
    lo, hi = 4.0, max(0.5 * L, 12.0)
    edges = np.linspace(lo, hi, n_windows + 1)
    w = edges[1] - edges[0]
    ranges = [(float(edges[k] - 0.25 * w), float(edges[k + 1] + 0.25 * w)) for k in range(n_windows)]
    names = [nm for nm, _ in ls.groups]
    gids = [names.index("end_a") + 1, names.index("end_b") + 1]
    r0 = end_to_end(ls)
    init = int(np.clip(np.searchsorted(edges, r0) - 1, 0, n_windows - 1))
    q = [((k,), 1.0) for k in range(n_windows)]
    return VcmdDef(interval=100, dims=[(gids, ["end_a", "end_b"], ranges)], init_vs=[init], seed=seed, q_cano=q)


# ---- the five workloads ----

def config0(seed: int = 0) -> Tuple[LaunchSet, dict]:
    """2,000 rigid waters in a 39.1 A box (LJ on O only), 12 A, ZD alpha 0, NVE + SHAKE, dt 2 fs."""
    ls = build_system(2000, seed=1000 + seed, L=39.1, h_lj=False)
    cfg = dict(mode="md", integrator="leapfrog-presto", cutoff=12.0, **{"nsgrid-cutoff": 13.0},
               **{"nsgrid-update-intvl": 10}, **{"n-steps": 1000}, **{"time-step": 2.0},
               electrostatic="zero-dipole", **{"ele-alpha": 0.0}, constraint="shake",
               **{"const-tolerance": 1e-8}, **{"const-max-loops": 1000}, thermostat="none", temperature=300.0,
               **{"com-motion": "none"}, **{"print-interval-log": 100}, **{"random-seed": seed})
    return ls, cfg


def config1(seed: int = 0) -> Tuple[LaunchSet, dict]:
    """33,000 waters + 200-bead chain (~99.6 A box), 12 A, Langevin 300 K; water
    hydrogens carry LJ (sigma 0.4 A, epsilon 0.046 kcal/mol) beside the charged chain."""
    ls = build_system(33000, seed=2000 + seed, n_beads=200)
    cfg = dict(mode="md", integrator="langevin", cutoff=12.0, **{"nsgrid-cutoff": 13.0},
               **{"nsgrid-update-intvl": 10}, **{"n-steps": 10000}, **{"time-step": 2.0},
               electrostatic="zero-dipole", **{"ele-alpha": 0.0}, constraint="shake-settle",
               **{"const-tolerance": 1e-8}, **{"const-max-loops": 1000}, thermostat="none", temperature=300.0,
               **{"langevin-gamma": 1.0}, **{"com-motion": "none"}, **{"print-interval-log": 1000},
               **{"random-seed": seed})
    return ls, cfg


def config2(seed: int = 0) -> Tuple[LaunchSet, dict]:
    """400,000 waters (~228.7 A box, LJ on O only), 10 A, ZD alpha 0.1."""
    ls = build_system(400000, seed=3000 + seed, h_lj=False)
    cfg = dict(mode="md", integrator="leapfrog-presto", cutoff=10.0, **{"nsgrid-cutoff": 11.0},
               **{"nsgrid-update-intvl": 10}, **{"n-steps": 100}, **{"time-step": 2.0},
               electrostatic="zero-dipole", **{"ele-alpha": 0.1}, constraint="shake-settle",
               **{"const-tolerance": 1e-8}, **{"const-max-loops": 1000}, thermostat="none", temperature=300.0,
               **{"com-motion": "none"}, **{"print-interval-log": 100}, **{"random-seed": seed})
    return ls, cfg


def walker_sizes(n_walkers: int = 64, seed: int = 3, lo: int = 3000, hi: int = 30000):
    rng = np.random.default_rng(40000 + seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n_walkers)).astype(int)


def walker_system(n_atoms: int, seed: int, T: float = 300.0, vcmd: bool = True) -> LaunchSet:
    n_w = max((n_atoms - 200) // 3, 1)
    return build_system(n_w, seed=seed, n_beads=200, T=T, with_vcmd=vcmd)


def cfg_text(cfg: dict) -> str:
    lines = []
    for k, v in cfg.items():
        lines.append("--%-26s %s" % (k, v))
    return "\n".join(lines) + "\n"
