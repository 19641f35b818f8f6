"""The ISO_C_BINDING path: a Fortran program (fortran/gh_fortran_check.f90,
built against fortran/gems_hip.f90) drives the library like a GEMS build and
must reproduce the oracle's forces and the Python binding's trajectory."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests import common as cm

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "fortran", "_build", "gh_fortran_check")


@pytest.mark.parametrize("prec", [0, 1])
def test_fortran_binding_matches_oracle_and_python(tmp_path, prec):
    if not os.path.exists(EXE):
        pytest.fail("fortran/_build/gh_fortran_check is not built (make -C fortran)")
    from gems_amd import hip
    f = cm.load("au864_pbc")
    fin, fout = tmp_path / "in.txt", tmp_path / "out.txt"
    with open(fin, "w") as fh:
        fh.write("%d\n%.17e %.17e %.17e\n" % ((len(f["x"]),) + tuple(f["L"])))
        for p in f["x"]:
            fh.write("%.17e %.17e %.17e\n" % tuple(p))
    r = subprocess.run([EXE, str(fin), str(fout), str(prec)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = np.array([float(t) for t in open(fout).read().split()])
    n = len(f["x"])
    etot = vals[0]
    F = vals[1:1 + 3 * n].reshape(n, 3)
    en = vals[1 + 3 * n:].reshape(2, 3)
    s = O.smatb_params(cm.au_params(), 1)
    box = O.make_box(f["L"], f["pbc"])
    nn, lst, _ = O.neighbours(f["x"], box, O.smatb_rcut(s) + 1.0)
    Fo, eo, epo, _ = O.smatb_force(f["x"], f["types"], box, s, nn, lst)
    tol = 1e-10 if prec == 0 else 1e-4
    assert cm.force_err(F, Fo) <= tol
    # the same calls through the Python binding give identical results
    with hip.Context(0, prec) as c:
        c.set_box(f["L"], (1, 1, 1))
        c.set_atoms(f["x"])
        c.set_masses([cm.MASS_AU])
        c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
        Fp, _, etp = c.compute()
        c.prng_lcg(4242)
        c.set_gdist(300.0)
        c.integrator_vv(0.002)
        enp = c.run(20, each=10)
    assert np.array_equal(F, Fp) and etot == etp
    assert np.array_equal(en, enp)
