"""Parity of the HIP path (through the C ABI) with the oracle and with the
reference binary's golden fixtures.

Bars:
  * neighbour sets: bit-exact (sorted rows, caller indices);
  * plain FP64 form: max|dF| / rms|F| <= 1e-10, per-atom and total energy
    within 1e-10 relative, trajectories against the reference's 16-digit
    energy columns within 1e-9 (chaos amplifies last-bit differences);
  * mixed precision: max|dF| / rms|F| <= 1e-5 (the north-star bar on the
    force), total energy within 1e-7 relative.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import common as cm

pytestmark = pytest.mark.gpu

hip = pytest.importorskip("gems_amd.hip")


def _ctx(precision=hip.GH_FP64):
    return hip.Context(0, precision)


def _setup(name, precision=hip.GH_FP64, skin=1.0):
    f = cm.load(name)
    c = _ctx(precision)
    m = cm.meta()
    if name.startswith("agcu"):
        first = m["agcu_first_symbol"]
        c.set_box(None)
        c.set_atoms(f["x"], f["types"])
        c.set_masses(cm.agcu_masses(first))
        c.set_smatb(2, *cm.matrices(cm.agcu_params(first), 2))
        s = O.smatb_params(cm.agcu_params(first), 2)
        box = O.make_box(None)
    elif name.startswith("lj"):
        c.set_box(f["L"], f["pbc"])
        c.set_atoms(f["x"], f["types"])
        c.set_masses([1.0])
        c.set_lj(1.0, 1.0, 2.5)
        s = None
        box = O.make_box(f["L"], f["pbc"])
        skin = 0.3
    else:
        c.set_box(f["L"], f["pbc"])
        c.set_atoms(f["x"], f["types"])
        c.set_masses([cm.MASS_AU])
        c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
        s = O.smatb_params(cm.au_params(), 1)
        box = O.make_box(f["L"], f["pbc"])
    c.set_skin(skin)
    return f, c, s, box, skin


def _oracle_forces(f, s, box, skin, lj=False):
    if lj:
        nn, lst, _ = O.neighbours(f["x"], box, 2.5 + skin)
        F, e, ep = O.lj_force(f["x"], box, 1.0, 1.0, 2.5, nn, lst)
    else:
        nn, lst, _ = O.neighbours(f["x"], box, O.smatb_rcut(s) + skin)
        F, e, ep, _ = O.smatb_force(f["x"], f["types"], box, s, nn, lst)
    return F, e, ep, nn, lst


NAMES = ["au256_pbc", "au864_pbc", "agcu201_free", "lj4000_pbc"]


@pytest.mark.parametrize("name", NAMES)
def test_neighbour_sets_bit_exact(name):
    f, c, s, box, skin = _setup(name)
    c.compute()
    gn, gl = c.neighbors(cap=256)
    _, _, _, nn, lst = _oracle_forces(f, s, box, skin, lj=name.startswith("lj"))
    assert np.array_equal(gn, nn)
    for i in range(len(nn)):
        assert np.array_equal(gl[i, :gn[i]], np.sort(lst[i, :nn[i]])), i
    st = c.stats()
    assert st["maxnn"] == nn.max()
    c.close()


@pytest.mark.parametrize("name", NAMES)
def test_fp64_forces_match_oracle_and_reference(name):
    f, c, s, box, skin = _setup(name)
    F, e, et = c.compute()
    Fo, eo, epo, _, _ = _oracle_forces(f, s, box, skin, lj=name.startswith("lj"))
    assert cm.force_err(F, Fo) <= 1e-10
    assert np.max(np.abs(e - eo) / np.abs(eo)) <= 1e-10
    assert et == pytest.approx(epo, rel=1e-12)
    # the reference binary's own output (13 significant digits)
    assert cm.force_err(F, f["F"]) <= 1e-10
    assert et * cm.UI_EV == pytest.approx(float(f["epot_ev"]), rel=1e-12)
    c.close()


@pytest.mark.parametrize("name", NAMES)
def test_mixed_forces_error_bound(name):
    f, c, s, box, skin = _setup(name, precision=hip.GH_MIXED)
    F, e, et = c.compute()
    Fo, eo, epo, _, _ = _oracle_forces(f, s, box, skin, lj=name.startswith("lj"))
    rms = np.sqrt(np.mean(Fo ** 2))
    err_rms = np.sqrt(np.mean((F - Fo) ** 2)) / rms
    err_max = np.max(np.abs(F - Fo)) / rms
    print("%s mixed: rms %.3e max %.3e dE %.3e" % (name, err_rms, err_max, abs(et / epo - 1)))
    assert err_rms <= 3e-6
    assert err_max <= 1e-5
    assert et == pytest.approx(epo, rel=1e-7)
    c.close()


def _md_case(name, integ):
    f, c, s, box, skin = _setup(name)
    m = cm.meta()
    n = len(f["x"])
    if name.startswith("lj"):
        T, dt = m["lj_nve"]["T"], m["lj_nve"]["dt"]
    elif name.startswith("agcu"):
        T, dt = 600.0, 0.002
    else:
        T, dt = 300.0, 0.002
    c.prng_lcg(int(f["seed"]))
    if integ == "vv":
        c.set_gdist(T)
        c.integrator_vv(dt)
    else:
        c.integrator_ermak(dt, T, 1.0, rng=hip.GH_RNG_LCG)
    en = c.run(int(f["nsteps"]), each=int(f["each"]))
    x, v, F, e = c.state()
    return f, c, en, x


@pytest.mark.parametrize("name,integ", [("au864_nve", "vv"), ("lj4000_nve", "vv"),
                                        ("au864_ermak", "ermak"), ("agcu201_ermak", "ermak")])
def test_fp64_trajectory_matches_reference(name, integ):
    f, c, en, x = _md_case(name, integ)
    ref = f["energy"]
    err = np.max(np.abs(en - ref) / np.abs(ref).max(axis=0))
    print(name, "max rel energy deviation", err)
    assert err <= 1e-9
    d = x - f["x_final"]
    if "L" in f:
        d -= f["L"] * np.round(d / f["L"])
    assert np.abs(d).max() <= 1e-8
    c.close()


def test_reference_kat_bulk_ag4000():
    """examples/SMA-TB/bulk.gms on the GPU, against ref/Energy.bulk.dat."""
    cur = [np.array(p, float) for p in [(0, 0, 0), (0.5, 0.5, 0), (0, 0.5, 0.5), (0.5, 0, 0.5)]]
    for t in [(1, 0, 0), (0, 1, 0), (0, 0, 1)]:
        t = np.array(t, float)
        cur = cur + [a + t * j for a in cur for j in range(1, 10)]
    fact = 2.89 / np.sqrt(0.5)
    x = (np.array(cur) + 1.0) * fact
    c = _ctx()
    c.set_box([12.0 * fact] * 3, (0, 0, 0))
    c.set_atoms(x, np.zeros(len(x), np.int32))
    c.set_masses([cm.MASS_AG])
    r2, r3 = np.sqrt(2.0), np.sqrt(3.0)
    c.set_smatb(1, 0.1031, 1.1899, 10.85, 3.18, 2.89, 2.89 * r2, 2.89 * r3)
    c.set_skin(1.0)
    c.prng_lcg(123456)
    c.integrator_ermak(float(np.float32(0.002)), 1500.0, 1.0, rng=hip.GH_RNG_LCG)
    en = c.run(220, each=20)
    ref = np.loadtxt(cm.GOLDEN + "/ref_Energy.bulk.dat")
    assert np.all(np.abs(en - ref) <= 5.01e-7 * np.abs(ref))
    assert c.stats()["list_mode"] == 1
    c.close()


def test_reference_kat_md_ag13():
    f = cm.load("ag13_lbfgs")
    c = _ctx()
    c.set_box(None)
    c.set_atoms(f["x"], np.zeros(13, np.int32))
    c.set_masses([cm.MASS_AG])
    r2, r3 = np.sqrt(2.0), np.sqrt(3.0)
    c.set_smatb(1, 0.1031, 1.1899, 10.85, 3.18, 2.89, 2.89 * r2, 2.89 * r3)
    c.prng_lcg(123456)
    c.integrator_ermak(0.001, 300.0, 1.0, rng=hip.GH_RNG_LCG)
    en = c.run(1000, each=100)
    ref = np.loadtxt(cm.GOLDEN + "/ref_Energy.md.dat")
    assert np.all(np.abs(en - ref) <= 5.01e-7 * np.abs(ref))
    c.close()


def test_tile_shapes_and_autotune():
    # brick shapes change only how cells are grouped per workgroup: the
    # neighbour sets, and the FP64 forces bit for bit, must not depend on them
    x, L = _au_big(nx=12, seed=3)
    s = O.smatb_params(cm.au_params(), 1)
    box = O.make_box([L] * 3, (1, 1, 1))
    nn, lst, _ = O.neighbours(x, box, O.smatb_rcut(s) + 1.0)
    Fo, _, epo, _ = O.smatb_force(x, np.zeros(len(x), np.int32), box, s, nn, lst)
    ref = {}
    for prec in (hip.GH_FP64, hip.GH_MIXED):
        for bd in [(1, 1, 1), (2, 2, 2), (4, 4, 3), (3, 2, 1), None]:
            c = _ctx(prec)
            c.set_box([L] * 3, (1, 1, 1))
            c.set_atoms(x)
            c.set_masses([cm.MASS_AU])
            c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
            c.set_tile(bd)
            F, e, et = c.compute()
            assert c.stats()["list_mode"] == 1
            gn, gl = c.neighbors(cap=256)
            assert np.array_equal(gn, nn)
            for i in range(0, len(nn), 97):
                assert np.array_equal(gl[i, :gn[i]], np.sort(lst[i, :nn[i]])), (bd, i)
            if prec == hip.GH_FP64:
                assert cm.force_err(F, Fo) <= 1e-10
            else:
                assert np.sqrt(np.mean((F - Fo) ** 2)) / np.sqrt(np.mean(Fo ** 2)) <= 1e-5
            if prec in ref:
                tol = 0.0 if prec == hip.GH_FP64 else 2e-5
                assert np.max(np.abs(F - ref[prec])) <= tol * np.abs(Fo).max(), bd
            else:
                ref[prec] = F
            if bd is None:
                shape, us = c.autotune(reps=2)
                assert all(1 <= v <= 4 for v in shape) and us > 0
                F2, _, _ = c.compute()
                tol = 0.0 if prec == hip.GH_FP64 else 2e-5
                assert np.max(np.abs(F2 - ref[prec])) <= tol * np.abs(Fo).max()
            c.close()


def test_results_are_deterministic():
    out = []
    for _ in range(2):
        f, c, s, box, skin = _setup("au864_pbc", precision=hip.GH_MIXED)
        c.prng_lcg(1)
        c.set_gdist(300.0)
        c.integrator_vv(0.002)
        c.run(30)
        out.append(c.state())
        c.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


# --------------------------------------------------------------------------
# full-size properties (1,000,188 atoms: 63^3 fcc cells, the scale workload)
# --------------------------------------------------------------------------
def _au_big(nx=63, jitter=0.02, seed=1):
    a = 4.078
    x = cm.fcc(nx, a)
    if jitter:
        x = x + np.random.default_rng(seed).normal(0, jitter, x.shape)
    L = nx * a
    return np.mod(x, L), L


def test_full_size_perfect_lattice_closed_form():
    x, L = _au_big(jitter=0.0)
    m = cm.meta()["au_prm"]
    a = 4.078
    r = np.array([a / np.sqrt(2)] * 12 + [a] * 6)
    e_atom = np.sum(m["A"] * np.exp(-m["p"] * (r / m["r0"] - 1))) - np.sqrt(
        np.sum(m["xi"] ** 2 * np.exp(-2 * m["q"] * (r / m["r0"] - 1))))
    for prec, tol in [(hip.GH_FP64, 1e-12), (hip.GH_MIXED, 2e-6)]:
        c = _ctx(prec)
        c.set_box([L] * 3, (1, 1, 1))
        c.set_atoms(x)
        c.set_masses([cm.MASS_AU])
        c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
        F, e, et = c.compute()
        assert len(x) == 1000188
        assert np.abs(F).max() <= 1e-6 * cm.EV_UI  # 1e-6 eV/A
        assert np.max(np.abs(e * cm.UI_EV / e_atom - 1)) <= tol
        assert et * cm.UI_EV / len(x) == pytest.approx(e_atom, rel=tol)
        c.close()


def test_full_size_mixed_vs_fp64_and_momentum():
    x, L = _au_big()
    res = {}
    for prec in (hip.GH_FP64, hip.GH_MIXED):
        c = _ctx(prec)
        c.set_box([L] * 3, (1, 1, 1))
        c.set_atoms(x)
        c.set_masses([cm.MASS_AU])
        c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
        res[prec] = c.compute()
        c.close()
    F64, e64, et64 = res[hip.GH_FP64]
    Fmx, emx, etmx = res[hip.GH_MIXED]
    rms = np.sqrt(np.mean(F64 ** 2))
    # Newton's third law: the total force vanishes to rounding
    assert np.abs(F64.sum(axis=0)).max() <= 1e-9 * rms * len(x)
    err_rms = np.sqrt(np.mean((Fmx - F64) ** 2)) / rms
    err_max = np.max(np.abs(Fmx - F64)) / rms
    print("1M mixed vs fp64: rms %.3e max %.3e dE %.3e" % (err_rms, err_max, etmx / et64 - 1))
    assert err_rms <= 3e-6 and err_max <= 1e-5
    assert etmx == pytest.approx(et64, rel=1e-7)
    # a sample of atoms against the CPU oracle
    s = O.smatb_params(cm.au_params(), 1)
    box = O.make_box([L] * 3, (1, 1, 1))
    sel = [123457]
    # a 28 A cube around one atom: atoms within 2 A of its centre have complete
    # neighbourhoods out to 2 rce (their neighbours' densities included)
    c0 = x[sel[0]]
    d = x - c0
    d -= L * np.round(d / L)
    near = np.where(np.all(np.abs(d) < 14.0, axis=1))[0]
    nn, lst = O.neighbours_brute(x[near], box, O.smatb_rcut(s) + 1.0, strict=True)
    Fo, eo, _, _ = O.smatb_force(x[near], np.zeros(len(near), np.int32), box, s, nn, lst)
    core = np.where(np.all(np.abs(d[near]) < 2.0, axis=1))[0]  # complete shells only
    assert len(core) > 0
    assert np.max(np.abs(F64[near[core]] - Fo[core])) <= 1e-10 * rms


def test_philox_langevin_reaches_target_temperature():
    x, L = _au_big(nx=10)
    c = _ctx(hip.GH_MIXED)
    c.set_box([L] * 3, (1, 1, 1))
    c.set_atoms(x)
    c.set_masses([cm.MASS_AU])
    c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
    c.integrator_ermak(0.002, 600.0, 5.0, rng=hip.GH_RNG_PHILOX, seed=1234)
    c.run(1500)
    en = c.run(500, each=50)
    T = 2 * en[:, 1] / (cm.KB_EV * 3 * len(x))
    assert abs(T.mean() / 600.0 - 1) < 0.03
    c.close()


def test_overflow_and_errors_are_reported():
    c = _ctx()
    with pytest.raises(hip.GhError):
        c.compute()  # no atoms
    f = cm.load("au256_pbc")
    c.set_box(f["L"], f["pbc"])
    c.set_atoms(f["x"])
    with pytest.raises(hip.GhError):
        c.compute()  # no potential
    with pytest.raises(hip.GhError):
        c.set_skin(-1.0)
    c.close()


@pytest.mark.parametrize("name", ["au864_pbc", "agcu201_free", "lj4000_pbc"])
@pytest.mark.parametrize("prec", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_host_positions_keep_the_lists(name, prec):
    """gh_set_positions (the host integrates, the plugin computes forces):
    small moves keep the list, a host-side wrap is no displacement, a move
    past half the skin rebuilds; forces match the oracle each time."""
    f, c, s, box, skin = _setup(name, precision=prec)
    lj = name.startswith("lj")
    tol = 1e-10 if prec == hip.GH_FP64 else 1e-5
    c.compute()
    nupd0 = c.stats()["nupd"]
    rng = np.random.default_rng(11)
    x = f["x"] + rng.normal(0.0, 0.05 * skin, f["x"].shape)  # max |dx| stays below skin/2
    g = dict(f)
    g["x"] = x
    for step in range(3):
        xs = x.copy()
        if step == 1 and "L" in f and np.any(f["pbc"]):  # the host wraps some atoms
            L = np.asarray(f["L"], float)
            xs[::7] = np.where(np.asarray(f["pbc"], bool), xs[::7] + L, xs[::7])
        c.set_positions(xs)
        F, e, et = c.compute()
        Fo, eo, epo, _, _ = _oracle_forces(g, s, box, skin, lj=lj)
        assert np.sqrt(np.mean((F - Fo) ** 2)) / np.sqrt(np.mean(Fo ** 2)) <= tol, step
        assert c.stats()["nupd"] == nupd0, step  # no rebuild: |dx| < skin/2
    x = x.copy()
    x[0] += 0.6 * skin  # past half the skin: the next force call rebuilds
    g["x"] = x
    c.set_positions(x)
    F, e, et = c.compute()
    Fo, eo, epo, _, _ = _oracle_forces(g, s, box, skin, lj=lj)
    assert np.sqrt(np.mean((F - Fo) ** 2)) / np.sqrt(np.mean(Fo ** 2)) <= tol
    assert c.stats()["nupd"] == nupd0 + 1
    c.close()


@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_lbfgs_reaches_the_reference_minimum(precision):
    """gh_lbfgs (lbfgs_minimizator, Quasi_Newton.f90:160-232) from the start
    of examples/SMA-TB/md.gms (Ag13, Baletto parameters) reaches the minimum
    the reference's own `lbfgs` found (tests/golden/ag13_lbfgs.npz): the same
    energy (1e-9 eV/atom FP64, 1e-6 mixed) and shape (sorted pair distances
    within 1e-4 A FP64, 1e-3 mixed)."""
    lines = open(cm.GOLDEN + "/ag13_coords.xyz").read().splitlines()[2:15]
    x0 = np.array([[float(t) for t in ln.split()[1:4]] for ln in lines])
    ref = cm.load("ag13_lbfgs")
    r0 = 2.89
    m = [np.array([[v]]) for v in (0.1031, 1.1899, 10.85, 3.18, r0, r0 * np.sqrt(2.0),
                                   r0 * np.sqrt(3.0))]
    with hip.Context(0, precision) as c:
        c.set_box(None)
        c.set_atoms(x0)
        c.set_masses([float(np.float32(107.87))])
        c.set_smatb(1, *m)
        eps = 1e-8 if precision == hip.GH_FP64 else 1e-6
        it, ep = c.lbfgs(m=5, eps=eps, max_iter=2000)
        x, _, F, _ = c.state()

    def dists(y):
        d = np.linalg.norm(y[:, None] - y[None], axis=-1)
        return np.sort(d[np.triu_indices(len(y), 1)])

    fp64 = precision == hip.GH_FP64
    print("lbfgs %s: %d iterations, epot %.12f eV (reference %.12f), max |F| %.2e eV/A"
          % ("fp64" if fp64 else "mixed", it, ep, float(ref["epot_ev"]),
             np.abs(F).max() * cm.UI_EV))
    assert it < 2000
    assert ep == pytest.approx(float(ref["epot_ev"]), abs=(1e-9 if fp64 else 1e-6) * 13)
    assert np.abs(dists(x) - dists(ref["x"])).max() < (1e-4 if fp64 else 1e-3)
