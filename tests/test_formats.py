"""GEMS's on-disk formats (gems_amd/formats.py) against files the reference
binary wrote for the NVE example (tests/golden/make_golden.py --plj):
`checkpoint write` at steps 100 and 200 (chp_plj_s*.chp, write_chp,
src/Checkpoint.F90:154-187), the position frame at step 100
(xyz_plj_s100.xyz, write_pos, src/Output.F90:900-918) and the default energy
columns every 50 steps (energy_plj.dat, write_energy, src/Output.F90:824-831).
The writers reproduce the reference's bytes (checkpoints, atom lines of xyz
frames, energy rows); the readers agree with each other across formats.
"""
import os
import tempfile

import numpy as np
import pytest

from gems_amd import formats as F
from tests import common as cm


def _g(name):
    return os.path.join(cm.GOLDEN, name)


def test_checkpoint_fields():
    ck = F.read_chp(_g("chp_plj_s100.chp"))
    assert ck["nlocal"] == 200 and ck["prng"] == "lcg"
    assert np.array_equal(ck["tbox"], np.diag([50.0, 50.0, 50.0]))
    assert ck["piston"] == (False, False, False)
    assert ck["time"] == pytest.approx(0.2, rel=1e-12)  # 100 steps of 2 fs
    assert ck["dm_steps"] == 100.0
    for nm in ("pos", "vel", "acel"):
        assert ck[nm].shape == (200, 3) and np.all(np.isfinite(ck[nm]))
    assert np.all((ck["pos"] >= 0.0) & (ck["pos"] < 50.0))
    # velocity Verlet leaves acel = F/m: zero mean force, so sum(m a) = 0
    f = np.load(_g("nve_plj.npz"))
    m = f["mass"][f["types"]]
    assert np.abs((m[:, None] * ck["acel"]).sum(axis=0)).max() < 1e-12 * np.abs(
        m[:, None] * ck["acel"]).max() * 200


def test_checkpoint_positions_match_the_xyz_frame():
    ck = F.read_chp(_g("chp_plj_s100.chp"))
    syms, x, nframe, time = F.read_xyz(_g("xyz_plj_s100.xyz"))[0]
    assert syms == ["He"] * 100 + ["Ne"] * 100
    assert time == pytest.approx(ck["time"], rel=1e-14)
    # e25.12: 12 significant digits
    assert np.max(np.abs(x - ck["pos"]) / np.maximum(np.abs(ck["pos"]), 1e-300)) < 5e-12


def test_writers_reproduce_the_reference_bytes():
    with tempfile.TemporaryDirectory() as d:
        for name in ("chp_plj_s100.chp", "chp_plj_s200.chp"):
            out = os.path.join(d, name)
            F.write_chp(out, F.read_chp(_g(name)))
            assert open(out, "rb").read() == open(_g(name), "rb").read()
        syms, x, nframe, time = F.read_xyz(_g("xyz_plj_s100.xyz"))[0]
        out = os.path.join(d, "x.xyz")
        F.write_xyz(out, syms, x, nframe, time)
        mine = open(out).read().splitlines()
        ref = open(_g("xyz_plj_s100.xyz")).read().splitlines()
        # the count and (nframe, time) lines are list-directed (their spacing
        # is the compiler's); the atom lines are e25.12 and must match exactly
        assert mine[2:] == ref[2:]
        assert int(mine[0]) == int(ref[0])
        E = F.read_energy(_g("energy_plj.dat"))
        out = os.path.join(d, "e.dat")
        F.write_energy(out, E)
        assert open(out).read() == open(_g("energy_plj.dat")).read()


def test_fortran_e_editing():
    assert F.fortran_e(0.0, 25, 12) == "    0.000000000000E+00"[-25:].rjust(25)
    assert F.fortran_e(-0.001523390, 20, 7) == "      -0.1523390E-02"
    assert F.fortran_e(7.756065, 20, 7) == "       0.7756065E+01"
    assert F.fortran_e(1.5e120, 12, 4) == "  0.1500+121"
    assert F.fortran_e(-2.5e-150, 12, 4) == " -0.2500-149"
    assert F.fortran_e(1.0, 5, 4) == "*****"


def test_corrupt_checkpoint_is_refused():
    with tempfile.TemporaryDirectory() as d:
        b = bytearray(open(_g("chp_plj_s100.chp"), "rb").read())
        p = os.path.join(d, "bad.chp")
        open(p, "wb").write(bytes(b[:-5]))
        with pytest.raises(ValueError):
            F.read_chp(p)
