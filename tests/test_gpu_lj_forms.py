"""The Lennard-Jones family on the device path: the reference's own
examples/LJ/{lj,slj,sm1}.gms known-answer files (ref/Energy.lj.dat,
ref/Energy.slj.dat, ref/Energy.sm1.dat, 7 digits; lj and slj: two H atoms, a 251-point distance scan with
`out state`, then 1000 velocity-Verlet steps from rest, energy every 10),
and slj / wca / sm1 forces (slj_interact, wca_interact, Pairs.F90:403-560;
sm1_interact, :925-990) on a
two-type cluster against the oracle.

Bars: FP64 every row of the example files within 0.6 of their last digit;
mixed every entry within 2e-6 eV, i.e. 2e-6 of the well depth e = 1 eV (the
FP32 pair terms; measured 1.2e-6); forces max|dF|/rms|F| <= 1e-10 (FP64) /
1e-5 (mixed).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import common as cm

pytestmark = pytest.mark.gpu

hip = pytest.importorskip("gems_amd.hip")

SIG = 0.890898718140339


def _example(kind, precision):
    c = hip.Context(0, precision)
    c.set_box([cm.LJ_EXAMPLE_BOX] * 3, (0, 0, 0))
    c.set_atoms(np.array([[1.0, 1.0, 1.0], [2.2, 1.0, 1.0]]))
    if kind == "slj":
        c.set_slj([[1.0]], [[SIG]], [[1.108683]])
        c.set_masses([cm.MASS_H])
    else:
        c.set_lj(1.0, SIG, 2.0)
        c.set_masses([3.0])
    c.set_skin(1.0)
    return c


@pytest.mark.parametrize("kind", ["lj", "slj"])
@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_lj_examples_known_answer(kind, precision):
    with _example(kind, precision) as c:
        def at(x):
            c.set_positions(x)
            _, _, et = c.compute()
            return [et * cm.UI_EV, 0.0, et * cm.UI_EV]
        rows, x = cm.lj_example_scan(at)
        assert c.stats()["list_mode"] == 0
    with _example(kind, precision) as c:
        c.set_atoms(x)
        c.set_velocities(np.zeros((2, 3)))
        c.integrator_vv(cm.LJ_EXAMPLE_DT)
        E = np.vstack([rows, c.run(1000, each=10)])
    kat = np.loadtxt(cm.GOLDEN + "/ref_Energy.%s.dat" % kind)
    u = cm.last_digit_units(E, kat)
    rel = np.abs(E - kat) / (np.abs(kat) + 1e-300)
    print("%s %s: max %.2f units of the last digit; max rel %.2e (nonzero entries); "
          "max abs scan %.2e eV, dynamics %.2e eV (total column %.2e)"
          % (kind, "fp64" if precision == hip.GH_FP64 else "mixed", u.max(),
             rel[kat != 0].max(), np.abs(E - kat)[:251].max(), np.abs(E - kat)[251:].max(),
             np.abs(E - kat)[251:, 2].max()))
    if precision == hip.GH_FP64:
        assert np.all(u <= 0.6)
    else:
        assert np.all(np.abs(E - kat) <= 2e-6)


@pytest.mark.parametrize("kind", ["slj", "wca", "sm1"])
@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_slj_wca_forces_match_oracle(kind, precision):
    # a jittered two-type simple-cubic cluster, 6^3 sites 1.05 apart: every
    # pair region (12-6, spline, cut) is populated
    rng = np.random.default_rng(17)
    g = np.arange(6) * 1.05
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    x = x + rng.normal(0, 0.04, x.shape)
    types = (np.arange(len(x)) % 2).astype(np.int32)
    eps, sig = np.array([[0.5, 0.7], [0.7, 0.9]]), np.array([[0.9, 1.0], [1.0, 1.1]])
    rs = np.array([[1.05, 1.15], [1.15, 1.25]])
    if kind == "sm1":  # one (r1, r2, e) for all pairs: 0.9 < r < 1.6 is the ramp
        p = O.sm1_params(0.9, 1.6, 0.3, nt=2)
    else:
        p = O.slj_params(eps, sig, rs) if kind == "slj" else O.wca_params(eps, sig)
    box = O.make_box(None)
    nn, lst, _ = O.neighbours(x, box, p.rc + 0.5)
    Fo, eo, epo = O.plj_force(x, types, box, p, nn, lst)
    for boxed in (False, True):  # O(N^2) list, then the tiled path in a large box
        with hip.Context(0, precision) as c:
            if boxed:
                c.set_box([40.0] * 3, (0, 0, 0))
            else:
                c.set_box(None)
            c.set_atoms(x + (10.0 if boxed else 0.0), types)
            if kind == "slj":
                c.set_slj(eps, sig, rs)
            elif kind == "sm1":
                c.set_sm1(0.9, 1.6, 0.3)
            else:
                c.set_wca(eps, sig)
            c.set_skin(0.5)
            F, e, et = c.compute()
            mode = c.stats()["list_mode"]
        tol = 1e-10 if precision == hip.GH_FP64 else 1e-5
        assert mode == (1 if boxed else 0)
        assert cm.force_err(F, Fo) <= tol, (boxed, cm.force_err(F, Fo))
        assert et == pytest.approx(epo, rel=1e-10 if precision == hip.GH_FP64 else 1e-5)


@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_sm1_example_known_answer(precision):
    """examples/LJ/sm1.gms (ref/Energy.sm1.dat): two H atoms from one site, a
    301-point scan, then 1000 velocity-Verlet steps with atom 2 at 10 A/ps,
    energy every step.  Bars as above, the mixed one 2e-8 eV (2e-6 of the
    step's depth e = 0.01 eV)."""
    def ctx():
        c = hip.Context(0, precision)
        c.set_box(None)
        c.set_atoms(np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 1.0]]))
        c.set_sm1(1.0, 2.0, 0.01)
        c.set_masses([cm.MASS_H])
        c.set_skin(1.0)
        return c
    with ctx() as c:
        def at(x):
            c.set_positions(x)
            _, _, et = c.compute()
            return [et * cm.UI_EV, 0.0, et * cm.UI_EV]
        rows, x, v = cm.sm1_example_scan(at)
    with ctx() as c:
        c.set_atoms(x)
        c.set_velocities(v)
        c.integrator_vv(cm.SM1_DT)
        E = np.vstack([rows, c.run(1000, each=1)])
    kat = np.loadtxt(cm.GOLDEN + "/ref_Energy.sm1.dat")
    u = cm.last_digit_units(E, kat)
    print("sm1 %s: max %.2f units of the last digit, max abs %.2e eV"
          % ("fp64" if precision == hip.GH_FP64 else "mixed", u.max(), np.abs(E - kat).max()))
    if precision == hip.GH_FP64:
        assert np.all(u <= 0.6)
    else:
        assert np.all(np.abs(E - kat) <= 2e-8)
