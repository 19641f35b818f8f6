"""Two-type SMA-TB in a periodic box: the tiled passes with per-type-pair
parameters (the multi-type kernels, types staged in LDS) against the oracle,
at a size that runs four lanes per atom (4,000 atoms) and at one that runs
one lane per atom (131,072 atoms).  Bars as in test_gpu_parity.py: FP64
forces and per-atom energies within 1e-10, mixed max|dF|/rms|F| <= 1e-5.

The alloy is an fcc Ag-Cu solid solution (a = 3.87 A, the mean of the two
lattice constants), 50:50 labels from a seeded generator, 0.05 A Gaussian
jitter; parameters are the reference's Ag/Cu/AgCu rows with internal types
in order of first appearance (Ag first).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import common as cm

pytestmark = pytest.mark.gpu

hip = pytest.importorskip("gems_amd.hip")


def _alloy(ncell, seed=21):
    a = 3.87
    x = cm.fcc(ncell, a)
    rng = np.random.default_rng(seed)
    x += rng.normal(0.0, 0.05, x.shape)
    t = (rng.random(len(x)) < 0.5).astype(np.int32)
    t[0] = 0  # Ag first: internal type 0
    L = np.array([ncell * a] * 3)
    return x, t, L


@pytest.mark.parametrize("ncell", [10, 32], ids=["4000-four-lanes", "131072-one-lane"])
def test_alloy_forces_match_oracle(ncell):
    x, t, L = _alloy(ncell)
    rows = cm.agcu_params("Ag")
    s = O.smatb_params(rows, 2)
    box = O.make_box(L, [1, 1, 1])
    skin = 1.0
    nn, lst, _ = O.neighbours(x, box, O.smatb_rcut(s) + skin)
    Fo, eo, epo, _ = O.smatb_force(x, t, box, s, nn, lst)
    rms = np.sqrt(np.mean(Fo ** 2))
    for prec in (hip.GH_FP64, hip.GH_MIXED):
        with hip.Context(0, prec) as c:
            c.set_box(L, [1, 1, 1])
            c.set_atoms(x, t)
            c.set_masses(cm.agcu_masses("Ag"))
            c.set_smatb(2, *cm.matrices(rows, 2))
            c.set_skin(skin)
            F, e, et = c.compute()
            assert c.stats()["list_mode"] == 1  # the tiled cell path
            if prec == hip.GH_FP64:
                assert cm.force_err(F, Fo) <= 1e-10
                assert np.max(np.abs(e - eo) / np.abs(eo)) <= 1e-10
                assert et == pytest.approx(epo, rel=1e-12)
            else:
                err_max = np.max(np.abs(F - Fo)) / rms
                print("alloy %d mixed max|dF|/rms %.3e" % (len(x), err_max))
                assert err_max <= 1e-5
                assert et == pytest.approx(epo, rel=1e-7)
