"""Pins the CPU oracle (oracle/gems_oracle.c) before it is trusted as the
checker of the HIP path.

Three kinds of pins:
  * the reference's own known-answer files (examples/SMA-TB/ref/*.dat) and
    golden fixtures written by the reference binary (tests/golden/);
  * mathematics: closed forms, F = -grad E by finite differences, momentum
    conservation, perfect-lattice symmetry, brute-force neighbour sets;
  * the reference's rebuild count (nupd_vlist), which pins test_update.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import common as cm


def _nb(x, box, rl):
    nn, lst, mode = O.neighbours(x, box, rl)
    return nn, lst, mode


# --------------------------------------------------------------------------
# constants
# --------------------------------------------------------------------------
def test_constants_match_reference_literals():
    c = O.constants()
    assert c["ev_ui"] == pytest.approx(cm.EV_UI, rel=0, abs=0)
    assert c["kB_ui"] == cm.KB_UI
    assert c["ev_ui"] == pytest.approx(9648.6071219474234, rel=1e-15)
    assert c["kB_ui"] == pytest.approx(0.83145762283562896, rel=1e-15)
    assert O.element_mass(79) == cm.MASS_AU != 196.97


# --------------------------------------------------------------------------
# RNG: the reference stream is pinned by its own KAT below (bulk.gms) and by
# the tempgdist fixtures; here the algebra of the generator.
# --------------------------------------------------------------------------
def test_lcg_uniform_range_and_determinism():
    a, b = O.LCG(123456), O.LCG(123456)
    u = np.array([a.ranu() for _ in range(20000)])
    assert np.all(u == np.array([b.ranu() for _ in range(20000)]))
    assert u.min() > 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01
    # am = (1 - 2^-24) * 2^-31 exactly (single-precision constant folding)
    am = (1 - 2.0 ** -24) * 2.0 ** -31
    k = np.round(u / am)
    assert np.all(am * k == u) and np.all(k % 2 == 1)


def test_gasdev_moments():
    r = O.LCG(7)
    g = np.array([r.gasdev() for _ in range(200000)])
    assert abs(g.mean()) < 0.01
    assert abs(g.std() - 1) < 0.01


# --------------------------------------------------------------------------
# neighbour lists
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name,rl,strict", [("au864_pbc", None, True),
                                            ("lj4000_pbc", 2.8, True),
                                            ("au256_pbc", None, False),
                                            ("agcu201_free", None, False)])
def test_neighbour_sets_equal_bruteforce(name, rl, strict):
    f = cm.load(name)
    x = f["x"]
    box = O.make_box(f["L"], f["pbc"]) if "L" in f else O.make_box(None)
    if rl is None:
        if name.startswith("agcu"):
            s = O.smatb_params(cm.agcu_params(cm.meta()["agcu_first_symbol"]), 2)
        else:
            s = O.smatb_params(cm.au_params(), 1)
        rl = O.smatb_rcut(s) + 1.0
    nn, lst, mode = _nb(x, box, rl)
    assert mode == (1 if strict else 0)
    bn, bl = O.neighbours_brute(x, box, rl, strict=strict)
    assert np.array_equal(nn, bn)
    for i in range(len(x)):
        assert np.array_equal(np.sort(lst[i, :nn[i]]), bl[i, :bn[i]])
    # full lists are symmetric
    pairs = {(i, j) for i in range(len(x)) for j in lst[i, :nn[i]]}
    assert all((j, i) in pairs for (i, j) in pairs)


def test_linked_cell_order_is_reference_chain_order():
    # within the first stencil cell (nabor 0), the chain runs from the last
    # inserted atom: neighbours of the same cell come in descending index
    f = cm.load("lj4000_pbc")
    box = O.make_box(f["L"], f["pbc"])
    nn, lst, mode = _nb(f["x"], box, 2.8)
    L = f["L"][0]
    nc = int(L / 2.8)
    cell = lambda p: tuple((p / (L / nc)).astype(int))
    i = 0
    same = [j for j in lst[i, :nn[i]] if cell(f["x"][j]) == cell(f["x"][i])]
    assert same == sorted(same, reverse=True)


# --------------------------------------------------------------------------
# SMA-TB
# --------------------------------------------------------------------------
def test_smatb_tail_coefficients_closed_form():
    # b-polynomial P(x) = b1 x^5 + b2 x^4 + b3 x^3, x = r - rce, must match
    # xi*exp(-q(r/r0-1)) in value, slope and curvature at rci (Logsdail 2012),
    # and the r-polynomial must match A*exp(-p(r/r0-1)) likewise.
    for (A, xi, p, q, r0, rci, rce) in [(0.2061, 1.79, 10.229, 4.036, 2.884, 4.078, 4.994509586),
                                        (0.0855, 1.224, 10.96, 2.278, 2.556, 3.6148, 4.4271)]:
        s = O.smatb_params([(0, 0, A, xi, p, q, r0, rci, rce)], 1)
        b1, b2, b3 = s.b1[0], s.b2[0], s.b3[0]
        r1, r2, r3 = s.r1[0], s.r2[0], s.r3[0]
        x = rci - rce
        P = lambda c1, c2, c3: (c1 * x ** 5 + c2 * x ** 4 + c3 * x ** 3,
                                5 * c1 * x ** 4 + 4 * c2 * x ** 3 + 3 * c3 * x ** 2,
                                20 * c1 * x ** 3 + 12 * c2 * x ** 2 + 6 * c3 * x)
        e = xi * np.exp(-q * (rci / r0 - 1))
        v, d1, d2 = P(b1, b2, b3)
        assert v == pytest.approx(e, rel=1e-12)
        assert d1 == pytest.approx(-q / r0 * e, rel=1e-11)
        assert d2 == pytest.approx((q / r0) ** 2 * e, rel=1e-10)
        e = A * np.exp(-p * (rci / r0 - 1))
        v, d1, d2 = P(r1, r2, r3)
        assert v == pytest.approx(e, rel=1e-12)
        assert d1 == pytest.approx(-p / r0 * e, rel=1e-11)
        assert d2 == pytest.approx((p / r0) ** 2 * e, rel=1e-10)


@pytest.mark.parametrize("name", ["au256_pbc", "au864_pbc", "agcu201_free"])
def test_smatb_forces_match_reference_binary(name):
    f = cm.load(name)
    if name.startswith("agcu"):
        first = cm.meta()["agcu_first_symbol"]
        s = O.smatb_params(cm.agcu_params(first), 2)
        box = O.make_box(None)
    else:
        s = O.smatb_params(cm.au_params(), 1)
        box = O.make_box(f["L"], f["pbc"])
    nn, lst, _ = _nb(f["x"], box, O.smatb_rcut(s) + 1.0)
    F, e, ep, _ = O.smatb_force(f["x"], f["types"], box, s, nn, lst)
    # fixtures hold 13 significant digits
    assert cm.force_err(F, f["F"]) < 1e-10
    assert np.max(np.abs(e * cm.UI_EV - f["e_atom_ev"]) / np.abs(f["e_atom_ev"])) < 1e-11
    assert ep * cm.UI_EV == pytest.approx(f["epot_ev"], rel=1e-13)


def test_lj_forces_match_reference_binary():
    f = cm.load("lj4000_pbc")
    box = O.make_box(f["L"], f["pbc"])
    nn, lst, _ = _nb(f["x"], box, 2.8)
    F, e, ep = O.lj_force(f["x"], box, 1.0, 1.0, 2.5, nn, lst)
    assert cm.force_err(F, f["F"]) < 1e-10
    assert ep * cm.UI_EV == pytest.approx(f["epot_ev"], rel=1e-13)


def _tb_energy(x, types, box, s):
    nn, lst, _ = _nb(x, box, O.smatb_rcut(s) + 1.0)
    F, e, ep, _ = O.smatb_force(x, types, box, s, nn, lst)
    return F, ep


def test_smatb_force_is_minus_gradient():
    # F = -dE/dx by central differences on a jittered Au cluster whose pairs
    # span both the exponential and the quintic-tail regions
    rng = np.random.default_rng(3)
    x = cm.fcc(3, 4.078)[:60] + rng.normal(0, 0.15, (60, 3))
    box = O.make_box(None)
    s = O.smatb_params(cm.au_params(), 1)
    types = np.zeros(60, np.int32)
    F, _ = _tb_energy(x, types, box, s)
    h = 1e-5
    for (i, d) in [(0, 0), (7, 1), (23, 2), (41, 0), (59, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[i, d] += h
        xm[i, d] -= h
        g = (_tb_energy(xp, types, box, s)[1] - _tb_energy(xm, types, box, s)[1]) / (2 * h)
        assert -g == pytest.approx(F[i, d], rel=2e-6, abs=1e-4)
    assert np.abs(F.sum(axis=0)).max() < 1e-9 * np.abs(F).max() * len(x)


def test_smatb_perfect_lattice_closed_form():
    # perfect fcc Au: forces vanish, and every atom has 12 neighbours at
    # a/sqrt2 (exponential), 6 at a (= rci, exponential branch of pass 2 but
    # the tail branch of pass 1 where the polynomial equals the exponential),
    # and 24 at a*sqrt(1.5) = rce (contributing zero).
    a = 4.078
    m = cm.meta()["au_prm"]
    x = cm.fcc(5, a)
    box = O.make_box([5 * a] * 3, (1, 1, 1))
    s = O.smatb_params(cm.au_params(), 1)
    F, ep = _tb_energy(x, np.zeros(len(x), np.int32), box, s)
    assert np.abs(F).max() < 1e-13 * cm.EV_UI  # 1e-13 eV/A: rounding only
    r = np.array([a / np.sqrt(2)] * 12 + [a] * 6)
    rep = np.sum(m["A"] * np.exp(-m["p"] * (r / m["r0"] - 1)))
    rho = np.sum(m["xi"] ** 2 * np.exp(-2 * m["q"] * (r / m["r0"] - 1)))
    e_atom = rep - np.sqrt(rho)
    assert ep / len(x) * cm.UI_EV == pytest.approx(e_atom, rel=1e-12)


# --------------------------------------------------------------------------
# trajectories: integrators + loop + test_update, against the reference
# --------------------------------------------------------------------------
def _md_from_fixture(name, nthreads=4):
    f = cm.load(name)
    m = cm.meta()
    if name.startswith("agcu"):
        first = m["agcu_first_symbol"]
        s = O.smatb_params(cm.agcu_params(first), 2)
        masses = np.array(cm.agcu_masses(first))[f["types"]]
        box = O.make_box(None)
        pot, lj = "smatb", None
    elif name.startswith("lj"):
        s = None
        masses = np.ones(len(f["x"]))
        box = O.make_box(f["L"], f["pbc"])
        pot, lj = "lj", (1.0, 1.0, 2.5)
    else:
        s = O.smatb_params(cm.au_params(), 1)
        masses = np.full(len(f["x"]), cm.MASS_AU)
        box = O.make_box(f["L"], f["pbc"])
        pot, lj = "smatb", None
    rng = O.LCG(int(f["seed"]))
    return f, s, masses, box, pot, lj, rng


@pytest.mark.parametrize("name,integ,T,skin,dt", [
    ("au864_nve", "vv", 300.0, 1.0, 0.002),
    ("lj4000_nve", "vv", None, 0.3, None),
    ("au864_ermak", "ermak", 300.0, 1.0, 0.002),
    ("agcu201_ermak", "ermak", 600.0, 1.0, 0.002)])
def test_md_matches_reference_binary(name, integ, T, skin, dt):
    f, s, masses, box, pot, lj, rng = _md_from_fixture(name)
    m = cm.meta()
    if name.startswith("lj"):
        T, dt = m["lj_nve"]["T"], m["lj_nve"]["dt"]
    v = np.zeros_like(f["x"])
    if integ == "vv":
        v = O.set_gdist(rng, T, masses)
    md = O.MDRun(f["x"], v, masses, box, pot=pot, tb=s, types=f["types"], lj=lj, skin=skin,
                 integ=integ, dt=dt, T=T, gamma=1.0, rng=rng, nthreads=4)
    en = md.run(int(f["nsteps"]), each=int(f["each"]))
    ref = f["energy"]
    # energy columns at 16 digits; chaos amplifies last-bit differences only mildly
    assert np.max(np.abs(en - ref) / np.abs(ref).max(axis=0)) < 1e-9
    # rebuild count (nupd_vlist) pins test_update, including the pos_old quirk
    assert md.nupd == int(f["nupd"])
    assert np.max(np.abs(md.x - f["x_final"])) < 1e-8


def test_reference_kat_bulk_ag4000_ermak():
    """examples/SMA-TB/bulk.gms against ref/Energy.bulk.dat (7 digits)."""
    cur = [np.array(p, float) for p in [(0, 0, 0), (0.5, 0.5, 0), (0, 0.5, 0.5), (0.5, 0, 0.5)]]
    for t in [(1, 0, 0), (0, 1, 0), (0, 0, 1)]:  # `reply t 10` (create_reply)
        t = np.array(t, float)
        cur = cur + [a + t * j for a in cur for j in range(1, 10)]
    fact = 2.89 / np.sqrt(0.5)
    x = (np.array(cur) + 1.0) * fact  # set move 1 1 1; set expand fact
    box = O.make_box([12.0 * fact] * 3, (0, 0, 0))
    s = O.smatb_params([(0, 0, 0.1031, 1.1899, 10.85, 3.18, 2.89, 2.89 * np.sqrt(2.0),
                         2.89 * np.sqrt(3.0))], 1)
    mass = np.full(len(x), O.element_mass(47))
    dt = float(np.float32(0.002))  # default dt is a default-real literal
    md = O.MDRun(x, np.zeros_like(x), mass, box, pot="smatb", tb=s, types=np.zeros(len(x)),
                 skin=1.0, integ="ermak", dt=dt, T=1500.0, gamma=1.0, rng=O.LCG(123456),
                 nthreads=4)
    en = md.run(220, each=20)
    ref = np.loadtxt(cm.GOLDEN + "/ref_Energy.bulk.dat")
    assert np.all(np.abs(en - ref) <= 5.01e-7 * np.abs(ref) + 1e-300)


def test_reference_kat_md_ag13_ermak():
    """examples/SMA-TB/md.gms (after lbfgs) against ref/Energy.md.dat."""
    f = cm.load("ag13_lbfgs")
    prm = [(0, 0, 0.1031, 1.1899, 10.85, 3.18, 2.89, 2.89 * np.sqrt(2.0), 2.89 * np.sqrt(3.0))]
    s = O.smatb_params(prm, 1)
    mass = np.full(13, O.element_mass(47))
    md = O.MDRun(f["x"], np.zeros((13, 3)), mass, O.make_box(None), pot="smatb", tb=s,
                 types=np.zeros(13), skin=1.0, integ="ermak", dt=0.001, T=300.0, gamma=1.0,
                 rng=O.LCG(123456))
    en = md.run(1000, each=100)
    ref = np.loadtxt(cm.GOLDEN + "/ref_Energy.md.dat")
    assert np.all(np.abs(en - ref) <= 5.01e-7 * np.abs(ref))
    cvs = np.loadtxt(cm.GOLDEN + "/ref_Energy.cvs.dat")
    assert f["epot_ev"] == pytest.approx(cvs[0, 0], rel=5e-7)


def test_vv_time_reversibility():
    f = cm.load("au256_pbc")
    s = O.smatb_params(cm.au_params(), 1)
    box = O.make_box(f["L"], f["pbc"])
    masses = np.full(256, cm.MASS_AU)
    v = O.set_gdist(O.LCG(5), 300.0, masses)
    md = O.MDRun(f["x"], v, masses, box, pot="smatb", tb=s, types=f["types"], integ="vv",
                 dt=0.002)
    md.run(50, each=0)
    md2 = O.MDRun(md.x, -md.v, masses, box, pot="smatb", tb=s, types=f["types"], integ="vv",
                  dt=0.002)
    md2.run(50, each=0)
    d = md2.x - f["x"]
    d -= f["L"] * np.round(d / f["L"])
    assert np.abs(d).max() < 1e-9


def test_set_gdist_temperature_and_momentum():
    masses = np.full(500, cm.MASS_AU)
    v = O.set_gdist(O.LCG(11), 300.0, masses)
    p = (masses[:, None] * v).sum(axis=0)
    assert np.abs(p).max() < 1e-9
    ke = O.ekin(v, masses)
    assert 2 * ke / (cm.KB_UI * 3 * 500) == pytest.approx(300.0, rel=1e-12)


def test_ermak_coefficients_closed_form():
    k = O.ermak_coeffs(0.002, 1.0, 300.0)
    g, dt = 1.0, 0.002
    c0 = np.exp(-g * dt)
    assert k["c0"] == c0
    assert k["c1"] == pytest.approx((1 - c0) / g, rel=1e-15)
    # variance of the position noise (Allen-Tildesley): dt/g (2 - (3-4c0+c0^2)/(g dt))
    assert k["sdr"] ** 2 == pytest.approx(dt / g * (2 - (3 - 4 * c0 + c0 ** 2) / (g * dt)),
                                          rel=1e-12)
    assert k["crv1"] ** 2 + k["crv2"] ** 2 == pytest.approx(1.0, rel=1e-14)


def test_fire_reaches_the_reference_lbfgs_minimum():
    """or_fire (Bitzek et al. 2006) from the start of examples/SMA-TB/md.gms
    reaches the minimum the reference's own L-BFGS found (the ag13_lbfgs
    fixture): same energy, same shape (sorted pair distances), and an
    independent force evaluation there is below the tolerance."""
    lines = open(cm.GOLDEN + "/ag13_coords.xyz").read().splitlines()[2:15]
    x0 = np.array([[float(t) for t in ln.split()[1:4]] for ln in lines])
    ref = cm.load("ag13_lbfgs")
    prm = [(0, 0, 0.1031, 1.1899, 10.85, 3.18, 2.89, 2.89 * np.sqrt(2.0), 2.89 * np.sqrt(3.0))]
    s = O.smatb_params(prm, 1)
    r = O.fire(x0, np.zeros(13), s, np.full(13, O.element_mass(47)), ftol=1e-5)
    assert r["converged"]
    assert r["epot"] == pytest.approx(float(ref["epot_ev"]), abs=1e-9 * 13)

    def dists(x):
        d = np.linalg.norm(x[:, None] - x[None], axis=-1)
        return np.sort(d[np.triu_indices(len(x), 1)])

    assert np.abs(dists(r["x"]) - dists(ref["x"])).max() < 1e-4
    nn, lst, _ = O.neighbours(r["x"], O.make_box(None), O.smatb_rcut(s) + 1.0)
    F, _, ep, _ = O.smatb_force(r["x"], np.zeros(13), O.make_box(None), s, nn, lst)
    assert np.sqrt((F ** 2).sum(axis=1)).max() * cm.UI_EV < 1e-5
    assert ep * cm.UI_EV == pytest.approx(r["epot"], rel=1e-13)


# --------------------------------------------------------------------------
# shifted Lennard-Jones (plj, Pairs.F90:282-370)
# --------------------------------------------------------------------------
def _plj_energy(x, types, box, p):
    nn, lst, _ = O.neighbours(x, box, p.rc + 1.0)
    F, _, ep = O.plj_force(x, types, box, p, nn, lst)
    return F, ep


def test_plj_pair_closed_form():
    # two atoms: p(r) = 4e((s/r)^12 - (s/r)^6) - prc - frc (r - rc) with
    # prc = u(rc) and frc = 24e(2 s^12/rc^13 - s^6/rc^7) = -u'(rc)
    # (Pairs.F90:282-291, 313-314), i.e. p = u(r) - u(rc) + u'(rc)(r - rc):
    # the energy vanishes at the cut, the force there is -2 u'(rc) (the
    # linear term's sign, reproduced as written)
    e, s, rc = 0.0037, 2.6, 10.0
    p = O.plj_params([[e]], [[s]], rc)
    box = O.make_box(None)
    types = np.zeros(2, np.int32)

    def u(r):
        return 4 * e * ((s / r) ** 12 - (s / r) ** 6)

    def du(r):
        return 4 * e * (-12 * s ** 12 / r ** 13 + 6 * s ** 6 / r ** 7)

    for r in (2.4, 3.0, 5.5, 9.999):
        x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        F, ep = _plj_energy(x, types, box, p)
        want = u(r) - u(rc) + du(rc) * (r - rc)
        assert ep * cm.UI_EV == pytest.approx(want, rel=1e-10, abs=1e-17)
        assert F[1, 0] * cm.UI_EV == pytest.approx(-(du(r) + du(rc)), rel=1e-10, abs=1e-15)
        assert F[0, 0] == pytest.approx(-F[1, 0], rel=1e-14)
    x = np.array([[0.0, 0.0, 0.0], [rc * (1 - 1e-9), 0.0, 0.0]])
    F, ep = _plj_energy(x, types, box, p)
    assert abs(ep) * cm.UI_EV < 3 * abs(du(rc)) * rc * 1e-9
    assert F[1, 0] * cm.UI_EV == pytest.approx(-2 * du(rc), rel=1e-6)


def test_plj_force_is_minus_gradient():
    # two types with their own (e, s); F = -dE/dx by central differences
    rng = np.random.default_rng(11)
    x = rng.uniform(0, 12.0, (40, 3))
    keep = [0]
    for i in range(1, len(x)):  # no pair closer than 2.3 A
        if np.min(np.linalg.norm(x[keep] - x[i], axis=1)) > 2.3:
            keep.append(i)
    x = x[keep]
    types = (np.arange(len(x)) % 2).astype(np.int32)
    p = O.plj_params([[0.002, 0.004], [0.004, 0.007]], [[2.2, 2.4], [2.4, 2.6]], 10.0)
    box = O.make_box(None)
    F, _ = _plj_energy(x, types, box, p)
    h = 1e-6
    for (i, d) in [(0, 0), (3, 1), (len(x) // 2, 2), (len(x) - 1, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[i, d] += h
        xm[i, d] -= h
        g = (_plj_energy(xp, types, box, p)[1] - _plj_energy(xm, types, box, p)[1]) / (2 * h)
        assert -g == pytest.approx(F[i, d], rel=1e-6, abs=1e-9)
    assert np.abs(F.sum(axis=0)).max() < 1e-10 * np.abs(F).max() * len(x)


def plj_nve_run(f, md_factory):
    """Energy rows at steps 100..5000 of the NVE example from its step-1
    state: 99 steps to the first row, then 4900 with a row every 100."""
    md = md_factory()
    return np.vstack([md.run(99, each=99), md.run(4900, each=100)])


def test_reference_kat_nve_plj():
    """examples/NVE/v_verlet.gms (100 He + 100 Ne, shifted LJ, VV) from the
    reference's state after step 1 (tests/golden/nve_plj.npz): the oracle
    against the reference's 16-digit columns of the same run and against the
    example's own ref/Energy.v_verlet.dat (7 digits).  The dilute gas is
    chaotic: the 16-digit rounding of the step-1 state grows to 1e-8 in epot
    by step 2000 and past the example's 7th digit after step 1800, so rows
    are compared in the window that rounding leaves; the total energy over
    all 50 rows."""
    f = np.load(cm.GOLDEN + "/nve_plj.npz")
    t = f["types"]
    box = O.make_box([float(f["L"])] * 3, (1, 1, 1))
    p = O.plj_params(f["eps"], f["sigma"], float(f["rc"]))
    E = plj_nve_run(f, lambda: O.MDRun(f["x1"], f["v1"], f["mass"][t], box, pot="plj", plj=p,
                                       types=t, skin=float(f["skin"]), integ="vv",
                                       dt=float(f["dt"])))
    ref = f["energy"]
    assert np.all(np.abs(E[:20, 0] - ref[:20, 0]) <= 1e-6 * np.abs(ref[:20, 0]))
    assert np.all(np.abs(E[:20, 1:] - ref[:20, 1:]) <= 1e-9 * np.abs(ref[:20, 1:]))
    kat = np.loadtxt(cm.GOLDEN + "/ref_Energy.v_verlet.dat")
    unit = 10.0 ** (np.floor(np.log10(np.abs(kat))) - 6)  # last printed digit
    assert np.all(np.abs(E[:18] - kat[:18]) <= 0.6 * unit[:18])
    assert np.all(np.abs(E[:, 2] - kat[:, 2]) <= 1e-5 * np.abs(kat[:, 2]))


# --------------------------------------------------------------------------
# splined LJ (slj) and WCA (Pairs.F90:374-560); the LJ examples' own files
# --------------------------------------------------------------------------
def _table_energy(p, box):
    def at(x):
        nn, lst, _ = O.neighbours(x, box, p.rc + 1.0)
        _, _, ep = O.plj_force(x, np.zeros(len(x), np.int32), box, p, nn, lst)
        return [ep * cm.UI_EV, 0.0, ep * cm.UI_EV]
    return at


def test_slj_and_wca_closed_forms():
    e, s = 0.7, 1.1
    box = O.make_box(None)

    def u(r):
        return 4 * e * ((s / r) ** 12 - (s / r) ** 6)

    def du(r):
        return 4 * e * (-12 * s ** 12 / r ** 13 + 6 * s ** 6 / r ** 7)

    # slj: value and slope continuous at rs, zero value and slope at rc
    rs = 1.3
    p = O.slj_params([[e]], [[s]], [[rs]])
    rc = p.rcp[0]
    assert rc == pytest.approx(rs - 32 * u(rs) / (11 * du(rs)), rel=1e-14)
    at = _table_energy(p, box)

    def pe(r):
        return at(np.array([[0.0, 0, 0], [r, 0, 0]]))[0]

    h = 1e-6
    assert pe(rs - 1e-12) == pytest.approx(pe(rs + 1e-12), rel=1e-9)
    assert (pe(rs - 1e-9) - pe(rs - 1e-9 - h)) / h == pytest.approx(
        (pe(rs + 1e-9 + h) - pe(rs + 1e-9)) / h, rel=1e-4)
    assert abs(pe(rc - 1e-7)) < 1e-20 and pe(rc + 1e-7) == 0.0
    assert pe(1.0) == pytest.approx(u(1.0), rel=1e-13)
    # wca: 12-6 + e below 2^(1/6) s (the exponent an FP32 literal), zero beyond
    p = O.wca_params([[e]], [[s]])
    at = _table_energy(p, box)
    rcw = 2.0 ** float(np.float32(1.0 / 6.0)) * s
    assert p.rcp[0] == rcw
    for r in (0.95, 1.0, 1.2):
        assert pe(r) == pytest.approx(u(r) + e, rel=1e-12)
    assert pe(rcw * (1 + 1e-9)) == 0.0 and abs(pe(rcw * (1 - 1e-12))) < 1e-9


@pytest.mark.parametrize("kind", ["slj", "wca"])
def test_slj_wca_force_is_minus_gradient(kind):
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 3.0, (24, 3))
    keep = [0]
    for i in range(1, len(x)):
        if np.min(np.linalg.norm(x[keep] - x[i], axis=1)) > 0.9:
            keep.append(i)
    x = x[keep]
    types = (np.arange(len(x)) % 2).astype(np.int32)
    eps, sig = [[0.5, 0.7], [0.7, 0.9]], [[0.9, 1.0], [1.0, 1.1]]
    p = (O.slj_params(eps, sig, [[1.05, 1.15], [1.15, 1.25]]) if kind == "slj"
         else O.wca_params(eps, sig))
    box = O.make_box(None)

    def energy(y):
        nn, lst, _ = O.neighbours(y, box, p.rc + 1.0)
        F, _, ep = O.plj_force(y, types, box, p, nn, lst)
        return F, ep

    F, _ = energy(x)
    h = 1e-6
    for (i, d) in [(0, 0), (1, 2), (len(x) - 1, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[i, d] += h
        xm[i, d] -= h
        g = (energy(xp)[1] - energy(xm)[1]) / (2 * h)
        assert -g == pytest.approx(F[i, d], rel=1e-5, abs=1e-7)


@pytest.mark.parametrize("kind", ["lj", "slj"])
def test_reference_kat_lj_examples(kind):
    """examples/LJ/{lj,slj}.gms against their ref/Energy.*.dat (7 digits):
    the distance scan and 1000 velocity-Verlet steps of two atoms."""
    box = O.make_box([cm.LJ_EXAMPLE_BOX] * 3, (0, 0, 0))
    if kind == "slj":
        p = O.slj_params([[1.0]], [[0.890898718140339]], [[1.108683]])
        at, mass = _table_energy(p, box), cm.MASS_H
        md_kw = dict(pot="plj", plj=p)
    else:
        def at(x):
            nn, lst, _ = O.neighbours(x, box, 3.0)
            _, _, ep = O.lj_force(x, box, 1.0, 0.890898718140339, 2.0, nn, lst)
            return [ep * cm.UI_EV, 0.0, ep * cm.UI_EV]
        mass = 3.0  # `set mass 3.`
        md_kw = dict(pot="lj", lj=(1.0, 0.890898718140339, 2.0))
    rows, x = cm.lj_example_scan(at)
    md = O.MDRun(x, np.zeros((2, 3)), np.full(2, mass), box, types=np.zeros(2), skin=1.0,
                 integ="vv", dt=cm.LJ_EXAMPLE_DT, **md_kw)
    E = np.vstack([rows, md.run(1000, each=10)])
    kat = np.loadtxt(cm.GOLDEN + "/ref_Energy.%s.dat" % kind)
    assert E.shape == kat.shape
    assert np.all(cm.last_digit_units(E, kat) <= 0.6)


# --------------------------------------------------------------------------
# ghost-image periodic boxes (box mic F)
# --------------------------------------------------------------------------
def test_ghost_images_lj_match_reference_binary():
    # 32 LJ atoms in a 3.36 sigma box (< 2 rc): the reference's ghost atoms
    # against a free-boundary supercell of every image within rc + skin
    f = cm.load("lj32_ghost")
    n = len(f["x"])
    X, _ = cm.ghost_supercell(f["x"], f["L"], 2.8)
    box = O.make_box(None)
    nn, lst, _ = O.neighbours(X, box, 2.8)
    F, e, ep = O.lj_force(X, box, 1.0, 1.0, 2.5, nn, lst)
    assert cm.force_err(F[:n], f["F"]) < 1e-10
    assert e[:n].sum() * cm.UI_EV == pytest.approx(f["epot_ev"], rel=1e-13)
    # the box is small enough that the nearest image alone is not the same
    boxp = O.make_box(f["L"], f["pbc"])
    nn, lst, _ = O.neighbours(f["x"], boxp, 2.8)
    Fm, _, epm = O.lj_force(f["x"], boxp, 1.0, 1.0, 2.5, nn, lst)
    assert cm.force_err(Fm, f["F"]) > 0.1
    assert abs(epm * cm.UI_EV - f["epot_ev"]) > 1.0


# --------------------------------------------------------------------------
# sm1 (Pairs.F90:916-990)
# --------------------------------------------------------------------------
def _sm1_pair(p, r):
    box = O.make_box(None)
    x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    nn, lst, _ = O.neighbours(x, box, p.rc + 0.5)
    F, _, ep = O.plj_force(x, np.zeros(2, np.int32), box, p, nn, lst)
    return F, ep * cm.UI_EV


def test_sm1_pair_closed_form():
    # -e inside r1, 0 at the cut, and the reference's force r dp/dr (its
    # factor r, reproduced as written) pulling the atoms together; the
    # derivative by central differences of the energy the oracle books
    r1, r2, e = 1.0, 2.0, 0.01
    p = O.sm1_params(r1, r2, e)
    for r in (0.5, 1.0):
        F, E = _sm1_pair(p, r)
        assert E == pytest.approx(-e, rel=1e-14) and np.all(F == 0.0)
    assert _sm1_pair(p, r2)[1] == pytest.approx(0.0, abs=1e-18)
    for r in (1.1, 1.37, 1.5, 1.83, 1.97):
        F, E = _sm1_pair(p, r)
        h = 1e-6
        dE = (_sm1_pair(p, r + h)[1] - _sm1_pair(p, r - h)[1]) / (2 * h)  # eV/A, pair total
        # the pair energy is split e/2 + e/2; atom 1 is pulled toward atom 2
        assert F[0, 0] * cm.UI_EV == pytest.approx(r * dE, rel=1e-7)
        assert F[1, 0] == pytest.approx(-F[0, 0], rel=1e-14)


def test_sm1_example_known_answer():
    # examples/LJ/sm1.gms against ref/Energy.sm1.dat (1301 rows, 7 digits)
    p = O.sm1_params(1.0, 2.0, 0.01)
    box = O.make_box(None)

    def at(x):
        nn, lst, _ = O.neighbours(x, box, p.rc + 1.0)
        _, _, ep = O.plj_force(x, np.zeros(2, np.int32), box, p, nn, lst)
        return [ep * cm.UI_EV, 0.0, ep * cm.UI_EV]
    rows, x, v = cm.sm1_example_scan(at)
    md = O.MDRun(x, v, np.full(2, cm.MASS_H), box, types=np.zeros(2), skin=1.0, integ="vv",
                 dt=cm.SM1_DT, pot="plj", plj=p)
    E = np.vstack([rows, md.run(1000, each=1)])
    kat = np.loadtxt(cm.GOLDEN + "/ref_Energy.sm1.dat")
    assert E.shape == kat.shape
    assert np.all(cm.last_digit_units(E, kat) <= 0.6)
