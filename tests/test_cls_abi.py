"""Launch-set (.cls) contract and the C-ABI library surface (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from omegagene_amd import cls, gen
from omegagene_amd import engine


def test_cls_roundtrip_and_section_sizes():
    ls = gen.build_system(300, seed=1, n_beads=20, with_vcmd=True)
    b = cls.write_cls(ls)
    sizes = cls.header_sizes(b)
    n = ls.n_atoms
    assert sizes[0] == 96                        # box: 12 doubles (Read.cpp:182-193)
    assert sizes[1] == sizes[2] == 4 + 24 * n    # crd / vel (Read.cpp:195-262)
    # n2 n3 n4 + trios {3 int, 3 double}; ala3 (9 pairs, 2047 trios, 9 quads) gives
    # 12 + 9*16 + 2047*36 + 9*64 = 74424 (samples/ala3/log03_md.txt:53)
    assert sizes[4] == 12 + 36 * len(ls.shake_trio_ids)
    r = cls.read_cls(b)
    assert np.array_equal(r.crd, ls.crd) and np.array_equal(r.vel, ls.vel)
    assert np.array_equal(r.atom_type, ls.atom_type)
    assert np.allclose(r.lj6, ls.lj6) and np.allclose(r.lj12, ls.lj12)
    assert np.array_equal(r.nb15off, ls.nb15off)
    assert np.array_equal(r.shake_trio_ids, ls.shake_trio_ids)
    assert np.allclose(r.shake_trio_d, ls.shake_trio_d)
    assert np.array_equal(r.bond_ids, ls.bond_ids)
    assert [g for g, _ in r.groups] == [g for g, _ in ls.groups]
    assert r.vcmd is not None and r.vcmd.dims[0][2] == ls.vcmd.dims[0][2]
    assert cls.write_cls(r) == b


def test_cls_rejects_other_versions():
    ls = gen.build_system(100, seed=2, L=15.0)
    b = bytearray(cls.write_cls(ls))
    b[8:14] = b"v.0.40"
    with pytest.raises(ValueError):
        cls.read_cls(bytes(b))


def test_nb15off_table_reference_layout():
    ls = gen.build_system(50, seed=3, L=12.0, n_beads=6)
    tab = ls.nb15off_table()
    w0 = tab[0]
    assert list(w0[:2]) == [1, 2] and w0[2] == -1
    assert tab.shape == (ls.n_atoms, 32)
    ex = ls.excess_pairs()
    assert all(a < b for a, b in ex)
    # beads: 1-2, 1-3, 1-4 exclusions
    b0 = 150
    assert set(tab[b0][tab[b0] >= 0]) == {b0 + 1, b0 + 2, b0 + 3}


def test_library_exports_every_declared_symbol():
    if not os.path.exists(engine.LIB_PATH):
        from omegagene_amd import build
        build.build()
    lib = ctypes.CDLL(engine.LIB_PATH)
    names = engine.exports()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_cites_reference_for_each_group():
    hdr = open(os.path.join(os.path.dirname(engine.LIB_PATH), "..", "include", "og_engine.h")).read()
    for sym in ("cuda_set_device", "cuda_alloc_atom_info", "cuda_pairwise_ljzd", "cuda_enumerate_cell_pairs",
                "cuda_memcpy_dtoh_work"):
        assert sym in hdr
    assert re.search(r"cuda_setup\.cu:\d+", hdr)


def test_params_from_cfg_units():
    p = engine.params_from_cfg({"langevin-gamma": 1.0, "integrator": "langevin", "constraint": "shake-settle",
                                "cutoff": 10.0})
    assert p.langevin_gamma == pytest.approx(1e-3)  # ps^-1 -> fs^-1 (Config.cpp:345-348)
    assert p.integrator == 1 and p.constraint == 2
    assert p.nsgrid_cutoff == 11.0                 # default cutoff + 1 when not given


def test_call_point_header_mirrors_reference_abi():
    """Every cuda_* call point of the reference's GPU path (declared at
    SubBox.cpp:5-78 and MiniCell.cpp:7-35) has an og_ counterpart in
    og_celeste.h that cites its definition, and the library exports it."""
    hdr = open(os.path.join(os.path.dirname(engine.LIB_PATH), "..", "include", "og_celeste.h")).read()
    points = ["hostalloc_atom_info", "hostfree_atom_info", "hostalloc_cell_info", "hostfree_cell_info",
              "hostalloc_atom_type_charge", "hostfree_atom_type_charge", "alloc_atom_info", "free_atom_info",
              "alloc_set_lj_params", "free_lj_params", "set_pbc", "set_constant", "zerodipole_constant",
              "memcpy_htod_atom_info", "set_cell_constant", "memcpy_htod_atomids", "set_atominfo",
              "memcpy_htod_crd", "set_crd", "enumerate_cell_pairs", "pairwise_ljzd", "pair_sync",
              "memcpy_dtoh_work"]
    declared = engine.exports("og_celeste.h")
    for p in points:
        assert "og_" + p in declared, p
        assert re.search(r"cuda_%s cuda_setup\.cu:\d+" % p, hdr), p
    # cuda_set_device / cuda_thread_sync map onto og_set_device / og_thread_sync (og_engine.h)
    assert {"og_set_device", "og_thread_sync"} <= set(engine.exports("og_engine.h"))
    lib = ctypes.CDLL(engine.LIB_PATH)
    assert all(hasattr(lib, "og_" + p) for p in points)


def test_call_point_prototypes_match_the_reference_declarations():
    """tests/abi_check.cpp static_asserts each reference call-point prototype
    (SubBox.cpp:5-78, MiniCell.cpp:7-35) with the og_ prefix and T*& as T**
    against include/og_celeste.h; it must compile."""
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    inc = os.path.join(os.path.dirname(here), "include")
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I" + inc, os.path.join(here, "abi_check.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
