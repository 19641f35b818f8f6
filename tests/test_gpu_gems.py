"""GEMS itself on the device path: the drop-in at the reference's own plugin
boundary.

fortran/gems.mk builds GEMS from its sources with SMA-TB's force call
(smatb_interact, src/TB.F90:457-586) handed to fortran/gems_tb_hip.f90 when
GEMS_BACKEND=hip.  GEMS keeps its script interpreter, its velocity-Verlet
integrator on the host, its outputs; every force evaluation goes through
gh_set_positions + gh_compute.  The script is the one that wrote the
au864_nve fixture with the reference binary (tests/golden/make_golden.py,
md_script): 864 Au atoms, periodic, `prng lcg` seed 4242, `set tempgdist
300`, `evolve v_verlet`, dt 2 fs, `dinamica 200`, energy every 10 steps.

Bars: FP64 energy columns within 1e-9 relative of the reference's 16-digit
columns and final positions within 1e-8 A; mixed precision within 1e-6
relative and 1e-4 A.
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


@pytest.mark.parametrize("precision,etol,xtol", [("fp64", 1e-9, 1e-8), ("mixed", 1e-6, 1e-4)])
def test_gems_script_on_the_device_path(precision, etol, xtol):
    f = cm.load("au864_nve")
    a = 4.078
    wd = tempfile.mkdtemp(prefix="gems_hip_")
    try:
        s = mg.md_script(wd, ["Au"] * len(f["x"]), f["x"], f["L"], (1, 1, 1),
                         [("Au", "Au", mg.AU, a, a * np.sqrt(1.5))], int(f["seed"]),
                         "> sys\nset tempgdist 300.\n", "evolve v_verlet\n", int(f["nsteps"]),
                         int(f["each"]), dt=0.002)
        with open(os.path.join(wd, "in.gms"), "w") as fh:
            fh.write(s)
        env = dict(os.environ, GEMS_BACKEND="hip", GEMS_HIP_PRECISION=precision)
        print("gems_hip run:", precision, wd, flush=True)
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
