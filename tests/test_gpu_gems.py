"""GEMS itself on the device path: the drop-in at the reference's own plugin
boundary.

fortran/gems.mk builds GEMS from its sources with the SMA-TB and LJ force
calls (smatb_interact, src/TB.F90:457-586; lj_interact, src/Pairs.F90:207-279)
handed to fortran/gems_plugin_hip.f90 when GEMS_BACKEND=hip.  GEMS keeps its
script interpreter, its velocity-Verlet integrator on the host, its outputs;
every force evaluation goes through gh_set_positions + gh_compute.  The
scripts are the ones that wrote the au864_nve and lj4000_nve fixtures with the
reference binary (tests/golden/make_golden.py, md_script): 864 Au atoms
(`prng lcg` 4242, `set tempgdist 300`, dt 2 fs, `dinamica 200`) and 4000 LJ
atoms (seed 777, T* 0.72, dt 0.005 tau, `dinamica 100`), periodic, velocity
Verlet, energy every 10 steps; and the 32-atom LJ box with ghost images
(`box mic F`, side 3.36 sigma, `dinamica 400`).

Bars: FP64 energy columns within 1e-9 relative of the reference's 16-digit
columns and final positions within 1e-8 A; mixed precision within 1e-6
relative and 1e-4 A (LJ liquid: 1e-5 and 1e-3 A, its trajectory is more
chaotic).
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from tests import common as cm
from tests.golden import make_golden as mg

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEMS_HIP = os.path.join(REPO, "fortran", "_build", "gems")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(GEMS_HIP),
                                 reason="fortran/_build/gems not built (needs the GEMS sources)")]


CASES = [("au864_nve", "fp64", 1e-9, 1e-8), ("au864_nve", "mixed", 1e-6, 1e-4),
         ("lj4000_nve", "fp64", 1e-9, 1e-8), ("lj4000_nve", "mixed", 1e-5, 1e-3),
         ("lj32_ghost_nve", "fp64", 1e-9, 1e-8), ("lj32_ghost_nve", "mixed", 1e-5, 1e-3)]


def _script(name, f, wd):
    if name == "au864_nve":
        a = 4.078
        return mg.md_script(wd, ["Au"] * len(f["x"]), f["x"], f["L"], (1, 1, 1),
                            [("Au", "Au", mg.AU, a, a * np.sqrt(1.5))], int(f["seed"]),
                            "> sys\nset tempgdist 300.\n", "evolve v_verlet\n",
                            int(f["nsteps"]), int(f["each"]), dt=0.002)
    T = 0.72 / mg.KB_EV
    return mg.md_script(wd, ["Ar"] * len(f["x"]), f["x"], f["L"], (1, 1, 1),
                        "nb_dcut = 0.3\ninteract 1 under lj 1. 1. 2.5\n", int(f["seed"]),
                        "> sys\nset tempgdist %s\n" % mg.fmt(T), "evolve v_verlet\n",
                        int(f["nsteps"]), int(f["each"]), mass=1.0,
                        dt=0.005 * np.sqrt(1.0 / mg.EV_UI),
                        extra="box mic F\n" if "ghost" in name else "")


@pytest.mark.parametrize("name,precision,etol,xtol", CASES)
def test_gems_script_on_the_device_path(name, precision, etol, xtol):
    f = cm.load(name)
    wd = tempfile.mkdtemp(prefix="gems_hip_")
    try:
        s = _script(name, f, wd)
        with open(os.path.join(wd, "in.gms"), "w") as fh:
            fh.write(s)
        env = dict(os.environ, GEMS_BACKEND="hip", GEMS_HIP_PRECISION=precision)
        print("gems_hip run:", name, precision, wd, flush=True)
        r = subprocess.run([GEMS_HIP, "in.gms"], cwd=wd, capture_output=True, text=True,
                           timeout=240, env=env)
        log = open(os.path.join(wd, "in.log")).read()
        assert r.returncode == 0 and "#-ERR->" not in log, log[-2000:]
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)
        X = mg.read_frames(os.path.join(wd, "X.xyz"))[-1]
        assert E.shape == f["energy"].shape
        rel = np.max(np.abs(E - f["energy"]) / np.abs(f["energy"]))
        dx = np.max(np.abs(X - f["x_final"]))
        print("energy columns max rel diff %.3e, final positions max diff %.3e A" % (rel, dx))
        assert rel <= etol
        assert dx <= xtol
    finally:
        shutil.rmtree(wd)
