"""Shifted Lennard-Jones (`plj`, plj_set / plj_interact, src/Pairs.F90:282-370)
on the device path, against the oracle and the reference's own NVE example.

The workload is examples/NVE/v_verlet.gms: 100 He + 100 Ne placed by the
reference's `fill` in a 50 A periodic box, one plj table over the three type
pairs (rc = 10 A, the script form's cut), velocity Verlet dt = 2 fs, skin
1 A.  tests/golden/nve_plj.npz holds the reference binary's state after
step 1 and its 16-digit energy columns every 100 steps to step 5000; the
example's own ref/Energy.v_verlet.dat (7 digits) is
tests/golden/ref_Energy.v_verlet.dat.  The dilute gas is chaotic (the
16-digit rounding of the step-1 state reaches the 7th digit of epot after
step 1800), so rows are compared in the window that rounding leaves and the
total energy over all 50 rows.

Bars: forces against the oracle, max|dF|/rms|F| <= 1e-10 (FP64) / 1e-5
(mixed).  The energy trace, FP64: epot within 1e-6 and ekin, total within
1e-9 relative of the reference's 16-digit columns to step 2000, every column
within 0.6 of the example file's last digit to step 1800; mixed: the total
within 1e-6 relative of the example file and epot within 1e-5 eV to step
1800.  Both: no drift, the median total of the last ten rows within 2e-5
relative of the first (a single row can fall inside a close collision, where
velocity Verlet's total swings by +-2.5 meV in either precision and in the
reference's own per-step trace, so rows past the chaos window are compared
through the median).
"""
import numpy as np
import pytest

from gems_amd import formats as FM
from oracle import oracle as O
from tests import common as cm

pytestmark = pytest.mark.gpu

hip = pytest.importorskip("gems_amd.hip")


def _fixture():
    return np.load(cm.GOLDEN + "/nve_plj.npz")


def _ctx(precision, f, x, box=True):
    c = hip.Context(0, precision)
    if box:
        c.set_box([float(f["L"])] * 3, (1, 1, 1))
    else:
        c.set_box(None)
    c.set_atoms(x, f["types"][:len(x)])
    c.set_masses(f["mass"])
    c.set_plj(f["eps"], f["sigma"], float(f["rc"]))
    c.set_skin(float(f["skin"]))
    return c


def _oracle_forces(f, x, box):
    p = O.plj_params(f["eps"], f["sigma"], float(f["rc"]))
    nn, lst, _ = O.neighbours(x, box, float(f["rc"]) + float(f["skin"]))
    return O.plj_force(x, f["types"][:len(x)], box, p, nn, lst)


@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_plj_forces_match_oracle_tiled(precision):
    f = _fixture()
    Fo, eo, epo = _oracle_forces(f, f["x1"], O.make_box([float(f["L"])] * 3, (1, 1, 1)))
    with _ctx(precision, f, f["x1"]) as c:
        F, e, et = c.compute()
        assert c.stats()["list_mode"] == 1
    tol = 1e-10 if precision == hip.GH_FP64 else 1e-5
    assert cm.force_err(F, Fo) <= tol
    assert et == pytest.approx(epo, rel=1e-10 if precision == hip.GH_FP64 else 1e-5)


@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_plj_forces_match_oracle_free_cluster(precision):
    # the atoms of a 24 A corner of the box as a free cluster: the O(N^2) list
    f = _fixture()
    x = f["x1"]
    sel = np.all(x < 24.0, axis=1)
    order = np.argsort(~sel, kind="stable")  # the selected atoms first, types kept aligned
    g = dict(f)
    g["types"] = f["types"][order]
    xs = x[order][:int(sel.sum())]
    Fo, eo, epo = _oracle_forces(g, xs, O.make_box(None))
    with _ctx(precision, g, xs, box=False) as c:
        F, e, et = c.compute()
        assert c.stats()["list_mode"] == 0
    tol = 1e-10 if precision == hip.GH_FP64 else 1e-5
    assert cm.force_err(F, Fo) <= tol


@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_plj_nve_known_answer(precision):
    f = _fixture()
    with _ctx(precision, f, f["x1"]) as c:
        c.set_velocities(f["v1"])
        c.integrator_vv(float(f["dt"]))
        E = np.vstack([c.run(99, each=99), c.run(4900, each=100)])
    ref = f["energy"]
    kat = np.loadtxt(cm.GOLDEN + "/ref_Energy.v_verlet.dat")
    unit = 10.0 ** (np.floor(np.log10(np.abs(kat))) - 6)
    print("%s: epot to step 2000 max rel %.2e; last-digit units to step 1800 max %.2f; "
          "total over 50 rows max rel %.2e"
          % ("fp64" if precision == hip.GH_FP64 else "mixed",
             np.max(np.abs(E[:20, 0] - ref[:20, 0]) / np.abs(ref[:20, 0])),
             np.max(np.abs(E[:18] - kat[:18]) / unit[:18]),
             np.max(np.abs(E[:, 2] - kat[:, 2]) / np.abs(kat[:, 2]))))
    print("total minus the example file's, ueV:", np.round((E[:, 2] - kat[:, 2]) * 1e6, 1))
    if precision == hip.GH_FP64:
        assert np.all(np.abs(E[:20, 0] - ref[:20, 0]) <= 1e-6 * np.abs(ref[:20, 0]))
        assert np.all(np.abs(E[:20, 1:] - ref[:20, 1:]) <= 1e-9 * np.abs(ref[:20, 1:]))
        assert np.all(np.abs(E[:18] - kat[:18]) <= 0.6 * unit[:18])
    else:
        assert np.all(np.abs(E[:18, 2] - kat[:18, 2]) <= 1e-6 * np.abs(kat[:18, 2]))
        assert np.all(np.abs(E[:18, 0] - kat[:18, 0]) <= 1e-5)
    assert abs(np.median(E[-10:, 2]) - kat[0, 2]) <= 2e-5 * abs(kat[0, 2])


@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_restart_from_the_reference_checkpoint(precision):
    """The reference's `checkpoint write` at step 100 of the NVE example
    (tests/golden/chp_plj_s100.chp) loaded into a context (positions,
    velocities, box, the lcg stream's state): 100 more steps reach the
    reference's own checkpoint at step 200 of the same run (positions within
    1e-9 A FP64 / 1e-4 A mixed, minimum image) and the energy row at step 200
    of its 16-digit trace; the context's checkpoint reads back as written."""
    f = _fixture()
    ck100 = FM.read_chp(cm.GOLDEN + "/chp_plj_s100.chp")
    ck200 = FM.read_chp(cm.GOLDEN + "/chp_plj_s200.chp")
    L = float(f["L"])
    with hip.Context(0, precision) as c:
        FM.load_checkpoint(c, ck100, f["types"])
        assert c.lcg_state() == ck100["lcg"]
        c.set_masses(f["mass"])
        c.set_plj(f["eps"], f["sigma"], float(f["rc"]))
        c.set_skin(float(f["skin"]))
        c.integrator_vv(float(f["dt"]))
        en = c.run(100, each=100)
        x, v, _, _ = c.state()
        ck = FM.checkpoint_of(c, [L] * 3, f["mass"][f["types"]], time=ck200["time"])
    d = x - ck200["pos"]
    d -= L * np.round(d / L)
    xtol, etol = (1e-9, 1e-9) if precision == hip.GH_FP64 else (1e-4, 1e-6)
    print("restart: max |dx| %.2e A, max |dv| %.2e A/ps, total energy rel diff %.2e"
          % (np.abs(d).max(), np.abs(v - ck200["vel"]).max(),
             abs(en[0, 2] / f["energy"][1, 2] - 1)))
    assert np.abs(d).max() <= xtol
    assert np.abs(v - ck200["vel"]).max() <= xtol * 100
    assert abs(en[0, 2] / f["energy"][1, 2] - 1) <= etol
    with __import__("tempfile").TemporaryDirectory() as tmp:
        path = tmp + "/mine.chp"
        FM.write_chp(path, ck)
        back = FM.read_chp(path)
    assert back["nlocal"] == 200 and np.array_equal(back["pos"], x)
    assert back["lcg"] == ck100["lcg"]  # velocity Verlet draws nothing


@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_nvt_ermak_example_known_answer(precision):
    """examples/NVT/ermak.gms (ref/Energy.ermak.dat, 7 digits): the NVE
    example's gas under `evolve ermak 300 10` for 10^4 steps, energy and
    temperature every 100, from the reference's checkpoint written just before
    `dinamica` (tests/golden/chp_nvt_s0.chp: positions, velocities and the lcg
    stream after `fill` and `set tempgdist`).  The Langevin noise comes from
    the same lcg stream, drawn in the reference's order.  Temperature column:
    ekin / (3/2 N kB).  Bars: FP64 every entry of all 100 rows within 0.6 of
    the file's last digit (gamma = 10/ps damps the trajectories' differences,
    so no window is needed); mixed every entry within 2e-6 of the thermal
    scale (the largest ekin, or the largest temperature for that column)."""
    f = _fixture()
    ck = FM.read_chp(cm.GOLDEN + "/chp_nvt_s0.chp")
    kat = np.loadtxt(cm.GOLDEN + "/ref_Energy.ermak.dat")
    with hip.Context(0, precision) as c:
        FM.load_checkpoint(c, ck, f["types"])
        c.set_masses(f["mass"])
        c.set_plj(f["eps"], f["sigma"], float(f["rc"]))
        c.set_skin(float(f["skin"]))
        c.integrator_ermak(0.002, 300.0, 10.0, rng=hip.GH_RNG_LCG)
        en = c.run(10000, each=100)
    E = np.column_stack([en, en[:, 1] / (1.5 * len(f["types"]) * cm.KB_EV)])
    u = cm.last_digit_units(E, kat)
    print("nvt ermak %s: max %.2f units of the last digit (rows 1-10: %.2f), max rel %.2e"
          % ("fp64" if precision == hip.GH_FP64 else "mixed", u.max(), u[:10].max(),
             (np.abs(E - kat) / np.abs(kat)).max()))
    scale = np.array([np.abs(kat[:, 1]).max()] * 3 + [np.abs(kat[:, 3]).max()])
    dev = np.abs(E - kat).max(axis=0) / scale
    print("per column max |dE| / thermal scale:", dev)
    assert E.shape == kat.shape
    if precision == hip.GH_FP64:
        assert np.all(u <= 0.6)
    else:
        assert np.all(dev <= 2e-6)
