"""The on-disk writers (og_io.cpp) against the reference's own output files
(tests/golden/wons_output.npz: a Presto trajectory, a GROMACS .trr and a
restart written by the reference all-pairs build).  The files are parsed
and written again from the parsed values: the bytes must come back the same,
except the restart's title record (the reference copies its version string
into an uninitialised 80-byte buffer).  No GPU is used."""
import os
import struct

import numpy as np

from helpers import load_npz
from omegagene_amd import output


def presto_frames(b):
    """Frames of a Presto trajectory (WriteTrr.cpp:179-254)."""
    o, out = 0, []
    while o < len(b):
        m0, step = struct.unpack_from("<ii", b, o)
        vals = struct.unpack_from("<10f", b, o + 8)
        m1 = struct.unpack_from("<i", b, o + 48)[0]
        rec = struct.unpack_from("<i", b, o + 52)[0]
        crd = np.frombuffer(b, "<f4", rec // 4, o + 56).reshape(-1, 3)
        out.append(dict(m0=m0, m1=m1, step=step, vals=vals, crd=crd))
        o += 56 + rec + 4
    return out


def restart_fields(b):
    """Fields of a restart file (WriteTrr.cpp:258-313): records title (80),
    atom counts (8), step + four doubles (36), crd and vel (24 n each)."""
    n = struct.unpack_from("<i", b, 92)[0]
    step = struct.unpack_from("<i", b, 108)[0]
    t, etot, ekin, epot = struct.unpack_from("<4d", b, 112)
    crd = np.frombuffer(b, "<f8", 3 * n, 152).reshape(n, 3)
    vel = np.frombuffer(b, "<f8", 3 * n, 152 + 24 * n + 8).reshape(n, 3)
    return dict(n=n, step=step, time=t, etot=etot, ekin=ekin, epot=epot, crd=crd, vel=vel)


def gromacs_frames(b):
    o, out = 0, []
    while o < len(b):
        hdr = struct.unpack_from("<3i12s13i", b, o)
        x_size, v_size, f_size, natoms, step = hdr[11], hdr[12], hdr[13], hdr[14], hdr[15]
        t = struct.unpack_from("<2d", b, o + 4 * 3 + 12 + 4 * 13)
        ob = o + 4 * 3 + 12 + 4 * 13 + 16
        box = np.frombuffer(b, "<f8", 9, ob)
        crd = np.frombuffer(b, "<f8", x_size // 8, ob + 72).reshape(-1, 3)
        size = ob + 72 + x_size + v_size + f_size - o
        out.append(dict(raw=b[o:o + size], step=step, time=t[0], box=box, crd=crd, natoms=natoms))
        o += size
    return out


def test_presto_trajectory_bytes_round_trip(tmp_path):
    d = load_npz("wons_output.npz")
    ref = bytes(d["trr_presto"])
    frames = presto_frames(ref)
    assert [f["step"] for f in frames] == [5, 10, 15, 20]
    assert all(f["m0"] == 44 and f["m1"] == 0 for f in frames)   # the reference closes the header with 0
    path = str(tmp_path / "t.trr")
    for k, f in enumerate(frames):
        time_, cpu, tot, kin, temp, pot, _rmsf, vdw, _hyd, _rmsd = f["vals"]
        output.write_presto_frame(path, k > 0, f["step"], time_, cpu, tot, kin, temp, pot, vdw,
                                  f["crd"].astype(np.float64))
    assert open(path, "rb").read() == ref


def test_restart_bytes_round_trip(tmp_path):
    d = load_npz("wons_output.npz")
    ref = bytes(d["restart"])
    r = restart_fields(ref)
    assert r["step"] == 20 and r["etot"] == r["ekin"] + r["epot"]
    path = str(tmp_path / "r.restart")
    output.write_restart(path, r["step"], r["time"], r["epot"], r["ekin"], r["crd"], r["vel"])
    mine = open(path, "rb").read()
    assert len(mine) == len(ref)
    assert mine[:4] == ref[:4] and mine[84:] == ref[84:]
    assert mine[4:4 + len(output.TITLE)] == output.TITLE.encode()


def test_gromacs_trajectory_layout(tmp_path):
    """GROMACS frames: every header byte and the box equal the reference's; the
    coordinates (stored in nm as x * 0.1) written from x = stored / 0.1 come
    back within one double rounding."""
    d = load_npz("wons_output.npz")
    ref = bytes(d["trr_gromacs"])
    frames = gromacs_frames(ref)
    assert [f["step"] for f in frames] == [5, 10, 15, 20]
    L = (float(np.frombuffer(ref, "<f8", 1, 92)[0]) * 10.0,) * 3
    path = str(tmp_path / "g.trr")
    for k, f in enumerate(frames):
        output.write_gromacs_frame(path, k > 0, f["step"], f["time"], L, crd=f["crd"] / 0.1)
    mine = gromacs_frames(open(path, "rb").read())
    for a, b in zip(mine, frames):
        hdr = 4 * 3 + 12 + 4 * 13 + 16 + 72
        assert a["raw"][:hdr] == b["raw"][:hdr]
        assert np.allclose(a["crd"], b["crd"], rtol=4e-16, atol=0)
