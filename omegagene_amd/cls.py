"""Reader and writer for the celeste launch set (``.cls``, format ``v.0.52``).

This is the ``--inp`` contract of the reference: the byte layout follows the
reader ``Read::load_launch_set`` (src/celeste/core/Read.cpp:50-114; header
:116-180, box :182-193, crd/vel :195-262, topology :264-459, constraints
:461-519, groups :563-592, VcMD :705-779, potential :780-786) and the writer
``dump_mdinput`` of toolkit/mdinput_generator.py:261-365.  Everything is
native-endian (little-endian here); atom types and group members are 1-based
on disk and 0-based in memory, like the reference reader.  Constraint
distances are stored unsquared on disk (the reader squares them,
Read.cpp:477-512).
"""
from __future__ import annotations

import io
import struct
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

MAGIC = 66261          # CelesteObject.cpp:28
LS_VERSION = "v.0.52"  # CelesteObject.cpp:29
MAX_N_NB15OFF = 32     # CelesteObject.h:26


def _i2(a):
    return np.asarray(a, np.int32).reshape(-1, 2)


@dataclass
class VcmdDef:
    """Extended-ensemble VcMD block (Read.cpp:705-779)."""
    interval: int
    # per dimension: (group ids into the group table (1-based as on disk), names, [(min, max)] per state)
    dims: List[Tuple[List[int], List[str], List[Tuple[float, float]]]]
    init_vs: List[int]                       # 0-based virtual state per dimension
    seed: int
    q_cano: List[Tuple[Tuple[int, ...], float]]   # (0-based state tuple, weight)


@dataclass
class VmcmdDef:
    """V-McMD / V-AUS virtual-state block (Read.cpp:521-561)."""
    interval: int
    temperature: float
    # per state: (order, lambda_low, lambda_high, prob_down, prob_up, [order + 1 polynomial coefficients],
    # alpha_low, alpha_high)
    states: List[Tuple[int, float, float, float, float, List[float], float, float]]
    init_vs: int                             # 0-based
    seed: int


@dataclass
class GroupCoordDef:
    """Group-coordinate block for restarting V-AUS / VcMD (Read.cpp:646-703):
    the enhanced atom groups and their unwrapped coordinates."""
    header: str
    aus_type: int
    group_ids: List[int]                     # indices into the group table, 1-based as on disk
    crd: List[np.ndarray]                    # per group float64[n_atoms_in_group, 3]


@dataclass
class LaunchSet:
    box: np.ndarray                      # float64[12] = {Lx,0,0, 0,Ly,0, 0,0,Lz, cx,cy,cz}
    crd: np.ndarray                      # float64[n,3]  (Angstrom)
    vel: np.ndarray                      # float64[n,3]  (Angstrom/fs)
    charge: np.ndarray                   # float64[n]
    mass: np.ndarray                     # float64[n]
    atom_type: np.ndarray                # int32[n], 0-based
    n_lj_types: int
    lj6: np.ndarray                      # float64[n_types, n_types] (symmetric)
    lj12: np.ndarray
    hps_cutoff: Optional[np.ndarray] = None  # float64[n_types, n_types]: HPS switch distance (the "nbpair hps" section)
    hps_lambda: Optional[np.ndarray] = None  # float64[n_types, n_types]: HPS hydrophobicity scale
    bond_ids: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    bond_k: np.ndarray = field(default_factory=lambda: np.zeros(0))
    bond_r0: np.ndarray = field(default_factory=lambda: np.zeros(0))
    angles: list = field(default_factory=list)
    torsions: list = field(default_factory=list)
    impros: list = field(default_factory=list)
    nb14: list = field(default_factory=list)
    nb15off: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))  # unordered pairs
    shake_pair_ids: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    shake_pair_d: np.ndarray = field(default_factory=lambda: np.zeros(0))
    shake_trio_ids: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))
    shake_trio_d: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))   # d12, d23, d31
    shake_quads: list = field(default_factory=list)
    settle_ids: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))
    settle_d: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    groups: List[Tuple[str, np.ndarray]] = field(default_factory=list)   # 0-based atom ids
    has_dist_rest: bool = True
    has_pos_rest: bool = False
    vcmd: Optional[VcmdDef] = None
    vmcmd: Optional[VmcmdDef] = None
    group_coord: Optional[GroupCoordDef] = None
    pote: Optional[float] = None

    @property
    def n_atoms(self) -> int:
        return int(self.crd.shape[0])

    def lj_tables_f32(self):
        """The float tables the pair code reads (MmSystem.cpp:313-319, real_pw)."""
        return self.lj6.astype(np.float32), self.lj12.astype(np.float32)

    def bonded_tables(self):
        """Angles, torsions, impropers and 1-4 pairs as flat arrays, with the
        parameters the reference's MmSystem holds: angle (ids [n,3], k, theta0),
        torsion / improper (ids [n,4], energy, overlaps, symmetry, phase), and
        1-4 (ids [n,2], types [n,2], coeff_vdw, coeff_ele).  Read.cpp:438 passes
        the file's two 1-4 coefficients to set_nb14_param (MmSystem.cpp:395-407)
        in the other order, so the LJ scaling is the file's second value and the
        Coulomb scaling the first, as here."""
        def tors(rows):
            if not rows:
                return (np.zeros((0, 4), np.int32), np.zeros(0), np.zeros(0, np.int32), np.zeros(0, np.int32),
                        np.zeros(0))
            a = np.array([r[:4] for r in rows], np.int32)
            return (a, np.array([r[4] for r in rows], float), np.array([r[5] for r in rows], np.int32),
                    np.array([r[6] for r in rows], np.int32), np.array([r[7] for r in rows], float))
        ang = (np.array([r[:3] for r in self.angles], np.int32).reshape(-1, 3),
               np.array([r[3] for r in self.angles], float), np.array([r[4] for r in self.angles], float))
        nb = (np.array([r[:2] for r in self.nb14], np.int32).reshape(-1, 2),
              np.array([r[2:4] for r in self.nb14], np.int32).reshape(-1, 2),
              np.array([r[5] for r in self.nb14], float), np.array([r[4] for r in self.nb14], float))
        return {"angle": ang, "torsion": tors(self.torsions), "impro": tors(self.impros), "nb14": nb}

    def shake_quad_tables(self):
        """Four-atom SHAKE groups: ids [n,4] and squared distances [n,6] of the
        pairs 01 12 20 30 23 13 (Read.cpp:499-516 squares them)."""
        if not self.shake_quads:
            return np.zeros((0, 4), np.int32), np.zeros((0, 6))
        ids = np.array([q[:4] for q in self.shake_quads], np.int32)
        d = np.array([q[4:10] for q in self.shake_quads], float)
        return ids, d * d

    def hps_tables_f32(self):
        """HPS cutoff and lambda tables as the pair code reads them (MmSystem.cpp:320-325, real_pw)."""
        nt = self.n_lj_types
        hc = self.hps_cutoff if self.hps_cutoff is not None else np.zeros((nt, nt))
        hl = self.hps_lambda if self.hps_lambda is not None else np.zeros((nt, nt))
        return hc.astype(np.float32), hl.astype(np.float32)

    def nb15off_table(self, max_n: int = MAX_N_NB15OFF) -> np.ndarray:
        """int32[n, 32] partner lists, -1 terminated, in input order (MmSystem.cpp:478-525)."""
        n = self.n_atoms
        pairs = _i2(self.nb15off)
        both = np.concatenate([pairs, pairs[:, ::-1]], 0) if len(pairs) else pairs
        order = np.arange(len(both))
        # reference order: for input pair k, a gets b then b gets a
        if len(pairs):
            order = np.empty(len(both), np.int64)
            order[0::2] = np.arange(len(pairs))
            order[1::2] = np.arange(len(pairs)) + len(pairs)
            both = both[order]
        tab = np.full((n, max_n), -1, np.int32)
        if len(both) == 0:
            return tab
        # drop exact duplicates keeping the first occurrence
        key = both[:, 0].astype(np.int64) * (n + 1) + both[:, 1]
        _, first = np.unique(key, return_index=True)
        keep = np.sort(first)
        both = both[keep]
        a = both[:, 0]
        srt = np.argsort(a, kind="stable")
        a_s, b_s = a[srt], both[srt, 1]
        start = np.searchsorted(a_s, np.arange(n))
        slot = np.arange(len(a_s)) - start[a_s]
        if slot.max() >= max_n:
            raise ValueError("more than %d 1-5 exclusions for an atom (MmSystem.cpp:484-488)" % max_n)
        tab[a_s, slot] = b_s
        return tab

    def excess_pairs(self) -> np.ndarray:
        """Excluded pairs i<j in the reference order (MmSystem.cpp:700-736)."""
        tab = self.nb15off_table()
        out = []
        for a in range(self.n_atoms):
            for b in tab[a]:
                if b == -1:
                    break
                if a < b:
                    out.append((a, b))
        return np.asarray(out, np.int32).reshape(-1, 2)

    def d_free(self, constraint: bool = True) -> int:
        """Degrees of freedom (Read.cpp:90-94,215; DynamicsMode.cpp:61-65)."""
        d = 3 * self.n_atoms - 3
        if self.has_pos_rest:
            d += 3
        if constraint:
            d -= len(self.shake_pair_ids) + 3 * len(self.shake_trio_ids) + 6 * len(self.shake_quads)
            # SETTLE waters are rigid too; the reference never subtracts them because its
            # SETTLE is a stub (ConstraintShake.cpp:284-286)
            d -= 3 * len(self.settle_ids)
        return d


def _p(fmt, *v):
    return struct.pack("<" + fmt, *v)


def _rec(dtype_fields, n):
    return np.zeros(n, dtype=np.dtype(dtype_fields))


def write_cls(ls: LaunchSet, path=None) -> bytes:
    """Serialise a LaunchSet (layout of Read.cpp:50-794); writes to ``path`` if given."""
    n = ls.n_atoms
    box = _p("12d", *[float(x) for x in ls.box])
    crd = _p("i", n) + np.ascontiguousarray(ls.crd, "<f8").tobytes()
    vel = _p("i", n) + np.ascontiguousarray(ls.vel, "<f8").tobytes()

    t = io.BytesIO()
    t.write(_p("i", n))
    t.write(np.ascontiguousarray(ls.charge, "<f8").tobytes())
    t.write(np.ascontiguousarray(ls.mass, "<f8").tobytes())
    t.write((np.asarray(ls.atom_type, "<i4") + 1).tobytes())
    nt = ls.n_lj_types
    iu = np.triu_indices(nt)
    lj = _rec([("a", "<i4"), ("b", "<i4"), ("c", "<f8"), ("d", "<f8")], len(iu[0]))
    lj["a"], lj["b"] = iu[0] + 1, iu[1] + 1
    lj["c"], lj["d"] = ls.lj6[iu], ls.lj12[iu]
    t.write(_p("iii", 8 + lj.nbytes, nt, len(lj)) + lj.tobytes())
    hp = lj.copy()
    hp["c"] = ls.hps_cutoff[iu] if ls.hps_cutoff is not None else 0.0
    hp["d"] = ls.hps_lambda[iu] if ls.hps_lambda is not None else 0.0
    t.write(_p("i", hp.nbytes) + hp.tobytes())

    bd = _rec([("a", "<i4"), ("b", "<i4"), ("k", "<f8"), ("r", "<f8")], len(ls.bond_ids))
    if len(bd):
        bd["a"], bd["b"] = _i2(ls.bond_ids)[:, 0], _i2(ls.bond_ids)[:, 1]
        bd["k"], bd["r"] = ls.bond_k, ls.bond_r0
    t.write(_p("ii", 4 + bd.nbytes, len(bd)) + bd.tobytes())

    def block(items, fmt):
        body = b"".join(_p(fmt, *it) for it in items)
        return _p("ii", 4 + len(body), len(items)) + body

    t.write(block(ls.angles, "iiidd"))
    t.write(block(ls.torsions, "iiiidiidi"))
    t.write(block(ls.impros, "iiiidiidi"))
    t.write(block([(a, b, t1 + 1, t2 + 1, cv, ce) for a, b, t1, t2, cv, ce in ls.nb14], "iiiidd"))
    pr = _i2(ls.nb15off)
    both = np.empty((2 * len(pr), 2), "<i4")
    both[0::2] = pr
    both[1::2] = pr[:, ::-1]
    t.write(_p("ii", 4 + both.nbytes, len(both)) + both.tobytes())
    topol = t.getvalue()

    def cst(pid, pd, tid, td, quads):
        b = _p("iii", len(pid), len(tid), len(quads))
        pp = _rec([("a", "<i4"), ("b", "<i4"), ("d", "<f8")], len(pid))
        if len(pp):
            pp["a"], pp["b"], pp["d"] = pid[:, 0], pid[:, 1], pd
        tt = _rec([("a", "<i4"), ("b", "<i4"), ("c", "<i4"), ("d1", "<f8"), ("d2", "<f8"), ("d3", "<f8")],
                  len(tid))
        if len(tt):
            tt["a"], tt["b"], tt["c"] = tid[:, 0], tid[:, 1], tid[:, 2]
            tt["d1"], tt["d2"], tt["d3"] = td[:, 0], td[:, 1], td[:, 2]
        return b + pp.tobytes() + tt.tobytes() + b"".join(_p("iiiidddddd", *q) for q in quads)

    has_shake = len(ls.shake_pair_ids) or len(ls.shake_trio_ids) or len(ls.shake_quads)
    shake = cst(_i2(ls.shake_pair_ids), np.asarray(ls.shake_pair_d),
                np.asarray(ls.shake_trio_ids, np.int32).reshape(-1, 3), np.asarray(ls.shake_trio_d).reshape(-1, 3),
                ls.shake_quads) if has_shake else b""
    settle = cst(np.zeros((0, 2), np.int32), np.zeros(0), np.asarray(ls.settle_ids, np.int32).reshape(-1, 3),
                 np.asarray(ls.settle_d).reshape(-1, 3), []) if len(ls.settle_ids) else b""

    g = io.BytesIO()
    if ls.groups:
        g.write(_p("i", len(ls.groups)))
        for name, atoms in ls.groups:
            nb = name.encode() + b"\0"
            g.write(_p("i", len(nb)) + nb + _p("i", len(atoms)))
        for name, atoms in ls.groups:
            g.write((np.asarray(atoms, "<i4") + 1).tobytes())
    groups = g.getvalue()
    dist = _p("i", 0) if ls.has_dist_rest else b""
    pos = _p("i", 0) if ls.has_pos_rest else b""

    v = b""
    if ls.vcmd is not None:
        vc = ls.vcmd
        vb = io.BytesIO()
        vb.write(_p("ii", vc.interval, len(vc.dims)))
        for gids, gnames, ranges in vc.dims:
            vb.write(_p("ii", len(ranges), len(gids)))
            for gid, nm in zip(gids, gnames):
                nb = nm.encode() + b"\0"
                vb.write(_p("ii", gid, len(nb)) + nb)
            for lo, hi in ranges:
                vb.write(_p("dd", lo, hi))
        for s in vc.init_vs:
            vb.write(_p("i", s + 1))
        vb.write(_p("ii", vc.seed, len(vc.q_cano)))
        for st, q in vc.q_cano:
            vb.write(b"".join(_p("i", s + 1) for s in st) + _p("d", q))
        v = vb.getvalue()
    pote = _p("d", ls.pote) if ls.pote is not None else b""
    ext = b""
    if ls.vmcmd is not None:
        mc = ls.vmcmd
        eb = io.BytesIO()
        eb.write(_p("iid", len(mc.states), mc.interval, mc.temperature))
        for order, lo, hi, pdn, pup, poly, alo, ahi in mc.states:
            eb.write(_p("i", order) + _p("dddd", lo, hi, pdn, pup))
            eb.write(_p("%dd" % (order + 1), *poly[:order + 1]))
            eb.write(_p("dd", alo, ahi))
        eb.write(_p("ii", mc.init_vs + 1, mc.seed))
        ext = eb.getvalue()
    gcs = b""
    if ls.group_coord is not None:
        gc = ls.group_coord
        gb = io.BytesIO()
        hb = gc.header.encode() + b"\0"
        gb.write(_p("i", len(hb)) + hb + _p("ii", gc.aus_type, len(gc.group_ids)))
        gb.write(_p("%di" % len(gc.group_ids), *gc.group_ids))
        gb.write(_p("%di" % len(gc.crd), *[len(c) for c in gc.crd]))
        for c in gc.crd:
            gb.write(np.ascontiguousarray(c, "<f8").tobytes())
        gcs = gb.getvalue()

    ver = LS_VERSION.encode() + b"\0"
    sizes = [len(box), len(crd), len(vel), len(topol), len(shake), len(settle), len(ext), len(groups), len(dist),
             len(pos), len(v), len(gcs), len(pote)]
    out = (_p("i", MAGIC) + _p("i", len(ver)) + ver + _p("13i", *sizes) + box + crd + vel + topol + shake
           + settle + ext + groups + dist + pos + v + gcs + pote)
    if path is not None:
        with open(path, "wb") as f:
            f.write(out)
    return out


class _Rd:
    def __init__(self, b: bytes):
        self.b = b
        self.o = 0

    def get(self, fmt):
        v = struct.unpack_from("<" + fmt, self.b, self.o)
        self.o += struct.calcsize("<" + fmt)
        return v

    def i(self):
        return self.get("i")[0]

    def rec(self, fields, n):
        dt = np.dtype(fields)
        a = np.frombuffer(self.b, dt, n, self.o).copy()
        self.o += dt.itemsize * n
        return a

    def arr(self, dt, n):
        a = np.frombuffer(self.b, dt, n, self.o).copy()
        self.o += a.nbytes
        return a

    def cstr(self, ln):
        s = self.b[self.o:self.o + ln].split(b"\0")[0].decode()
        self.o += ln
        return s


def read_cls(src) -> LaunchSet:
    """Parse a v.0.52 launch set (Read.cpp:50-794). ``src``: path or bytes."""
    b = src if isinstance(src, (bytes, bytearray)) else open(src, "rb").read()
    r = _Rd(bytes(b))
    if r.i() != MAGIC:
        raise ValueError("bad magic number (Read.cpp:119-132)")
    ver = r.cstr(r.i())
    if ver != LS_VERSION:
        raise ValueError("launch-set version %r, %r required (Read.cpp:140-147)" % (ver, LS_VERSION))
    (s_box, s_crd, s_vel, s_top, s_cst, s_set, s_ext, s_grp, s_dr, s_pr, s_vc, s_gc, s_pote) = r.get("13i")
    box = np.array(r.get("12d")) if s_box else np.zeros(12)
    crd = vel = None
    if s_crd:
        n = r.i()
        crd = r.arr("<f8", 3 * n).reshape(n, 3)
    if s_vel:
        n = r.i()
        vel = r.arr("<f8", 3 * n).reshape(n, 3)
    if crd is None:
        raise ValueError("launch set without coordinates")
    n = crd.shape[0]
    ls = LaunchSet(box=box, crd=crd, vel=vel if vel is not None else np.zeros_like(crd), charge=np.zeros(n),
                   mass=np.zeros(n), atom_type=np.zeros(n, np.int32), n_lj_types=0, lj6=np.zeros((0, 0)),
                   lj12=np.zeros((0, 0)), has_dist_rest=False)
    if s_top:
        end = r.o + s_top
        if r.i() != n:
            raise ValueError("topology atom count mismatch (Read.cpp:272-276)")
        ls.charge = r.arr("<f8", n)
        ls.mass = r.arr("<f8", n)
        ls.atom_type = (r.arr("<i4", n) - 1).astype(np.int32)
        r.i()
        nt = r.i()
        npairs = r.i()
        lj = r.rec([("a", "<i4"), ("b", "<i4"), ("c", "<f8"), ("d", "<f8")], npairs)
        ls.n_lj_types = nt
        ls.lj6 = np.zeros((nt, nt))
        ls.lj12 = np.zeros((nt, nt))
        ls.lj6[lj["a"] - 1, lj["b"] - 1] = lj["c"]
        ls.lj6[lj["b"] - 1, lj["a"] - 1] = lj["c"]
        ls.lj12[lj["a"] - 1, lj["b"] - 1] = lj["d"]
        ls.lj12[lj["b"] - 1, lj["a"] - 1] = lj["d"]
        # "nbpair hps": cutoff and lambda per type pair (Read.cpp:322-334)
        r.i()
        hp = r.rec([("a", "<i4"), ("b", "<i4"), ("c", "<f8"), ("d", "<f8")], npairs)
        ls.hps_cutoff = np.zeros((nt, nt))
        ls.hps_lambda = np.zeros((nt, nt))
        ls.hps_cutoff[hp["a"] - 1, hp["b"] - 1] = hp["c"]
        ls.hps_cutoff[hp["b"] - 1, hp["a"] - 1] = hp["c"]
        ls.hps_lambda[hp["a"] - 1, hp["b"] - 1] = hp["d"]
        ls.hps_lambda[hp["b"] - 1, hp["a"] - 1] = hp["d"]
        r.i()
        nb = r.i()
        bd = r.rec([("a", "<i4"), ("b", "<i4"), ("k", "<f8"), ("r", "<f8")], nb)
        ls.bond_ids = np.stack([bd["a"], bd["b"]], 1).astype(np.int32)
        ls.bond_k, ls.bond_r0 = bd["k"].copy(), bd["r"].copy()

        def block(fmt):
            r.i()
            k = r.i()
            return [r.get(fmt) for _ in range(k)]

        ls.angles = block("iiidd")
        ls.torsions = block("iiiidiidi")
        ls.impros = block("iiiidiidi")
        ls.nb14 = [(a, b_, t1 - 1, t2 - 1, cv, ce) for a, b_, t1, t2, cv, ce in block("iiiidd")]
        r.i()
        k = r.i()
        pairs = r.arr("<i4", 2 * k).reshape(k, 2)
        lo = np.minimum(pairs[:, 0], pairs[:, 1]).astype(np.int64)
        hi = np.maximum(pairs[:, 0], pairs[:, 1]).astype(np.int64)
        _, first = np.unique(lo * (n + 1) + hi, return_index=True)
        ls.nb15off = pairs[np.sort(first)].astype(np.int32)
        r.o = end

    def cst():
        n2, n3, n4 = r.get("3i")
        pp = r.rec([("a", "<i4"), ("b", "<i4"), ("d", "<f8")], n2)
        tt = r.rec([("a", "<i4"), ("b", "<i4"), ("c", "<i4"), ("d1", "<f8"), ("d2", "<f8"), ("d3", "<f8")], n3)
        q = [r.get("iiiidddddd") for _ in range(n4)]
        return (np.stack([pp["a"], pp["b"]], 1).astype(np.int32), pp["d"].copy(),
                np.stack([tt["a"], tt["b"], tt["c"]], 1).astype(np.int32),
                np.stack([tt["d1"], tt["d2"], tt["d3"]], 1), q)

    if s_cst:
        ls.shake_pair_ids, ls.shake_pair_d, ls.shake_trio_ids, ls.shake_trio_d, ls.shake_quads = cst()
    if s_set:
        _, _, ls.settle_ids, ls.settle_d, _ = cst()
    if s_ext:
        # V-McMD / V-AUS block (Read.cpp:521-561)
        end = r.o + s_ext
        n_vs, interval = r.get("ii")
        temperature = r.get("d")[0]
        states = []
        for _ in range(n_vs):
            order = r.i()
            lo, hi, pdn, pup = r.get("dddd")
            poly = list(r.get("%dd" % (order + 1)))
            alo, ahi = r.get("dd")
            states.append((order, lo, hi, pdn, pup, poly, alo, ahi))
        init, seed = r.get("ii")
        ls.vmcmd = VmcmdDef(interval, temperature, states, init - 1, seed)
        r.o = end
    if s_grp:
        ng = r.i()
        names, counts = [], []
        for _ in range(ng):
            names.append(r.cstr(r.i()))
            counts.append(r.i())
        for nm, c in zip(names, counts):
            ls.groups.append((nm, (r.arr("<i4", c) - 1).astype(np.int32)))
    if s_dr:
        ls.has_dist_rest = True
        r.o += s_dr
    if s_pr:
        ls.has_pos_rest = True
        r.o += s_pr
    if s_vc:
        interval, dim = r.get("ii")
        dims = []
        for _ in range(dim):
            n_vs, n_grp = r.get("ii")
            gids, gnames = [], []
            for _ in range(n_grp):
                gids.append(r.i())
                gnames.append(r.cstr(r.i()))
            dims.append((gids, gnames, [r.get("dd") for _ in range(n_vs)]))
        init_vs = [r.i() - 1 for _ in range(dim)]
        seed, n_states = r.get("ii")
        q = []
        for _ in range(n_states):
            st = tuple(x - 1 for x in r.get("%di" % dim))
            q.append((st, r.get("d")[0]))
        ls.vcmd = VcmdDef(interval, dims, init_vs, seed, q)
    if s_gc:
        # group coordinates of the enhanced groups (Read.cpp:646-703)
        end = r.o + s_gc
        header = r.cstr(r.i())
        aus_type, ng = r.get("ii")
        gids = list(r.get("%di" % ng))
        counts = list(r.get("%di" % ng))
        ls.group_coord = GroupCoordDef(header, aus_type, gids, [r.arr("<f8", 3 * c).reshape(c, 3) for c in counts])
        r.o = end
    if s_pote:
        ls.pote = r.get("d")[0]
    return ls


def header_sizes(src) -> Tuple[int, ...]:
    """The 13 section sizes of the header (Read.cpp:149-161)."""
    b = src if isinstance(src, (bytes, bytearray)) else open(src, "rb").read(256)
    ln = struct.unpack_from("<i", b, 4)[0]
    return struct.unpack_from("<13i", b, 8 + ln)
