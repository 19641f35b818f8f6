"""The tiled passes at both lane counts against the oracle.

* Two-type SMA-TB in a periodic box (the multi-type kernels, per-type-pair
  parameters, types staged in LDS) at 4,000 atoms (four lanes per atom) and
  131,072 atoms (one lane per atom): an fcc Ag-Cu solid solution (a = 3.87 A,
  the mean of the two lattice constants), 50:50 labels from a seeded
  generator, 0.05 A Gaussian jitter; the reference's Ag/Cu/AgCu rows with
  internal types in order of first appearance (Ag first).
* LJ at 131,072 atoms (one lane per atom; the 4,000-atom fixtures of
  test_gpu_parity.py run four).

Bars as in test_gpu_parity.py: FP64 forces (and per-atom energies) within
1e-10, mixed max|dF|/rms|F| <= 1e-5.
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


def test_lj_one_lane_size_matches_oracle():
    # the LJ passes at one lane per atom (131,072 atoms; the 4,000-atom
    # fixtures of test_gpu_parity.py run four): fcc 32^3 at rho* 0.8442 with
    # 0.05 sigma jitter, rc 2.5 sigma, skin 0.3 sigma
    a = (4.0 / 0.8442) ** (1.0 / 3.0)
    x = cm.fcc(32, a)
    x += np.random.default_rng(9).normal(0.0, 0.05, x.shape)
    L = np.array([32 * a] * 3)
    box = O.make_box(L, [1, 1, 1])
    nn, lst, _ = O.neighbours(x, box, 2.5 + 0.3)
    Fo, eo, epo = O.lj_force(x, box, 1.0, 1.0, 2.5, nn, lst)
    rms = np.sqrt(np.mean(Fo ** 2))
    for prec in (hip.GH_FP64, hip.GH_MIXED):
        with hip.Context(0, prec) as c:
            c.set_box(L, [1, 1, 1])
            c.set_atoms(x, np.zeros(len(x), np.int32))
            c.set_masses([1.0])
            c.set_lj(1.0, 1.0, 2.5)
            c.set_skin(0.3)
            F, e, et = c.compute()
            assert c.stats()["list_mode"] == 1
            if prec == hip.GH_FP64:
                assert cm.force_err(F, Fo) <= 1e-10
                assert et == pytest.approx(epo, rel=1e-12)
            else:
                err_max = np.max(np.abs(F - Fo)) / rms
                print("lj %d mixed max|dF|/rms %.3e" % (len(x), err_max))
                assert err_max <= 1e-5
                assert et == pytest.approx(epo, rel=1e-7)
