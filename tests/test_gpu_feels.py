"""One-sided interactions (`interact A feels B`, Interaction.f90:69-77 and
128-131: a cross group with half and newton off) on the device path through
gh_set_pair_masks, against the reference's own known-answer files:

* examples/LJ/wca.gms (ref/Energy.wca.dat): two H atoms, the second feels the
  first through wca (e 1, s 0.8909); two `out state` rows, a 250-point
  distance scan, then 1000 velocity-Verlet steps, energy every 10;
* examples/LJ/feels.gms (ref/Energy.feels.dat): five H atoms, `interact 2 with
  2 slj` over atoms 1-4 and `interact 4 feels 3 slj` (atom 5 feels atoms 2
  and 4) in one context, as two mask bits; `out state`, then 1000 steps.

Bars: FP64 every entry within 0.6 of the files' last (7th) digit; mixed
within 2e-6 eV (2e-6 of the well depth) for wca.gms and 3e-6 eV for
feels.gms, whose five atoms collide repeatedly over the 1000 steps (measured
2.0e-6 eV at its largest).
"""
import numpy as np
import pytest

from tests import common as cm

pytestmark = pytest.mark.gpu

hip = pytest.importorskip("gems_amd.hip")

SIG = 0.890898718140339


def _wca_scan(energy_at):
    """wca.gms's `set move` additions replayed in order: the rows of its
    `out state` commands and the positions the dynamics starts from."""
    x = np.array([[1.0, 1.0, 1.0], [2.2, 1.0, 1.0]])
    rows = [energy_at(x)]
    x[1, 0] += -1.2
    x[1, 0] += SIG
    rows.append(energy_at(x))
    x[1, 0] += -SIG
    x[1, 0] += 0.9
    for _ in range(250):
        x[1, 0] += 0.01
        rows.append(energy_at(x))
    x[1, 0] += -2.5
    x[1, 0] += -0.9
    x[1, 0] += 0.95
    return np.array(rows), x


def _check(E, kat, precision, label, mixed_bar=2e-6):
    u = cm.last_digit_units(E, kat)
    print("%s %s: max %.2f units of the last digit, max abs %.2e eV"
          % (label, "fp64" if precision == hip.GH_FP64 else "mixed", u.max(),
             np.abs(E - kat).max()))
    if precision == hip.GH_FP64:
        assert np.all(u <= 0.6)
    else:
        assert np.all(np.abs(E - kat) <= mixed_bar)


def _ctx(precision, x):
    c = hip.Context(0, precision)
    c.set_box([cm.LJ_EXAMPLE_BOX] * 3, (0, 0, 0))
    c.set_atoms(x)
    c.set_masses([cm.MASS_H])
    c.set_skin(1.0)
    return c


@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_wca_feels_example_known_answer(precision):
    x0 = np.array([[1.0, 1.0, 1.0], [2.2, 1.0, 1.0]])
    with _ctx(precision, x0) as c:
        c.set_wca([[1.0]], [[SIG]])
        c.set_pair_masks([0, 1], [1, 0])  # atom 2 feels atom 1

        def at(x):
            c.set_positions(x)
            F, _, et = c.compute()
            assert np.all(F[0] == 0.0)  # no reaction on the felt atom
            return [et * cm.UI_EV, 0.0, et * cm.UI_EV]
        rows, x = _wca_scan(at)
        assert c.stats()["list_mode"] == 0
    with _ctx(precision, x) as c:
        c.set_wca([[1.0]], [[SIG]])
        c.set_pair_masks([0, 1], [1, 0])
        c.set_velocities(np.zeros((2, 3)))
        c.integrator_vv(cm.LJ_EXAMPLE_DT)
        E = np.vstack([rows, c.run(1000, each=10)])
        xf, _, _, _ = c.state()
        assert np.array_equal(xf[0], x[0])  # the felt atom never moves
    _check(E, np.loadtxt(cm.GOLDEN + "/ref_Energy.wca.dat"), precision, "wca.gms")


@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_feels_example_known_answer(precision):
    x = np.array([[1.0, 1.0, 1.0], [2.2, 1.0, 1.0], [1.0, 1.0, 5.0], [2.2, 1.0, 5.0],
                  [3.4, 1.0, 1.0]])
    with _ctx(precision, x) as c:
        c.set_slj([[1.0]], [[SIG]], [[1.108683]])
        # bit 1: `interact 2 with 2` over atoms 1-4; bit 2: atom 5 feels atoms 2, 4
        c.set_pair_masks([1, 1, 1, 1, 2], [1, 3, 1, 3, 0])
        c.set_velocities(np.zeros((5, 3)))
        c.integrator_vv(cm.LJ_EXAMPLE_DT)
        ep, ek = c.energies()
        E = np.vstack([[ep, ek, ep + ek], c.run(1000, each=10)])
    _check(E, np.loadtxt(cm.GOLDEN + "/ref_Energy.feels.dat"), precision, "feels.gms", 3e-6)


def test_masks_need_pair_potentials():
    with _ctx(hip.GH_FP64, np.array([[1.0, 1.0, 1.0], [3.5, 1.0, 1.0]])) as c:
        c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
        c.set_pair_masks([1, 1], [1, 1])
        with pytest.raises(RuntimeError):
            c.compute()
