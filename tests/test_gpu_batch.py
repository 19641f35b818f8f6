"""Batched free clusters (gh_batch_*): forces against the oracle per cluster,
Langevin equal to the single-system path on the same Philox stream, replica
exchange bookkeeping, FIRE relaxations against the oracle's FP64 FIRE."""
import numpy as np
import pytest

from gems_amd import clusters as CL
from oracle import oracle as O
from tests import common as cm

pytestmark = pytest.mark.gpu
hip = pytest.importorskip("gems_amd.hip")


def _agcu_small(ngeom=12, rungs=3, seed=7, nmax=400):
    return CL.agcu_re_batch(n_geom=ngeom, rungs=rungs, nmin=13, nmax=nmax, seed=seed)


def _oracle_agcu():
    rows = []
    m = CL.agcu_params()
    for i in range(2):
        for j in range(i, 2):
            rows.append((i, j) + tuple(mm[i, j] for mm in m))
    return O.smatb_params(rows, 2)


def _batch_agcu(prec, off, x, t):
    b = hip.Batch(0, prec)
    b.set_systems(off, x, t)
    b.set_smatb(2, *CL.agcu_params())
    b.set_masses([CL.MASS_AG, CL.MASS_CU])
    return b


@pytest.mark.parametrize("prec,tol,nmax", [(hip.GH_FP64, 1e-10, 400), (hip.GH_MIXED, 1e-5, 400),
                                           (hip.GH_FP64, 1e-10, 1200), (hip.GH_MIXED, 1e-5, 1200)])
def test_batch_forces_match_oracle(prec, tol, nmax):
    # nmax 1200: batches with a system over 512 atoms run 512-thread workgroups
    off, x, t, T, lad = _agcu_small(nmax=nmax, ngeom=8)
    with _batch_agcu(prec, off, x, t) as b:
        F, ep = b.compute()
    s = _oracle_agcu()
    box = O.make_box(None)
    for k in range(len(off) - 1):
        a0, a1 = off[k], off[k + 1]
        nn, lst, mode = O.neighbours(x[a0:a1], box, O.smatb_rcut(s) + 1.0)
        Fo, eo, epo, _ = O.smatb_force(x[a0:a1], t[a0:a1], box, s, nn, lst)
        assert cm.force_err(F[a0:a1], Fo) <= tol, k
        assert ep[k] == pytest.approx(epo * cm.UI_EV, rel=1e-10 if prec == 0 else 1e-6)


def test_batch_langevin_equals_single_system_path():
    # one cluster: the batch kernel and the single-system context draw the same
    # Philox normals (seed, step, atom) and do the same FP64 arithmetic
    off, x, t, T, lad = _agcu_small(ngeom=1, rungs=1)
    with _batch_agcu(hip.GH_FP64, off, x, t) as b:
        b.set_temperatures([600.0])
        ep, ek = b.langevin(200, 0.002, 1.0, 99)
        xb, vb = b.state()
    with hip.Context(0, hip.GH_FP64) as c:
        c.set_box(None)
        c.set_atoms(x, t)
        c.set_masses([CL.MASS_AG, CL.MASS_CU])
        c.set_smatb(2, *CL.agcu_params())
        c.integrator_ermak(0.002, 600.0, 1.0, rng=hip.GH_RNG_PHILOX, seed=99)
        c.run(200)
        xc, vc, _, _ = c.state()
        epc, ekc = c.energies()
    assert np.max(np.abs(xb - xc)) < 1e-9
    assert np.max(np.abs(vb - vc)) < 1e-7
    assert ep[0] == pytest.approx(epc, rel=1e-11)
    assert ek[0] == pytest.approx(ekc, rel=1e-9)


def test_batch_systems_are_independent():
    # a system's forces and its FIRE path do not depend on the other systems in
    # the batch: cluster k alone reproduces its slice of the batch bit for bit
    off, x, t, T, lad = _agcu_small()
    k = 3
    a0, a1 = off[k], off[k + 1]
    one = np.array([0, a1 - a0], np.int32)
    for prec in (hip.GH_FP64, hip.GH_MIXED):
        with _batch_agcu(prec, off, x, t) as b:
            F, ep = b.compute()
            r = b.fire(max_iter=300)
            xf, _ = b.state()
        with _batch_agcu(prec, one, x[a0:a1], t[a0:a1]) as b:
            F1, ep1 = b.compute()
            r1 = b.fire(max_iter=300)
            xf1, _ = b.state()
        assert np.array_equal(F[a0:a1], F1)
        assert ep[k] == ep1[0]
        assert np.array_equal(xf[a0:a1], xf1)
        assert r["iters"][k] == r1["iters"][0]


def test_langevin_ladder_temperatures():
    off, x, t, T, lad = _agcu_small(ngeom=6, rungs=4)
    with _batch_agcu(hip.GH_MIXED, off, x, t) as b:
        b.set_temperatures(T)
        b.langevin(3000, 0.002, 5.0, 11)
        eks = []
        for _ in range(5):
            ep, ek = b.langevin(200, 0.002, 5.0, 11)
            eks.append(ek)
    n = np.diff(off)
    Tm = 2 * np.mean(eks, axis=0) / (3 * n * cm.KB_EV)
    # KE about the CM of free clusters: 3n - 3 degrees of freedom thermalised
    Tm *= 3 * n / (3 * n - 3)
    big = n > 100
    assert np.all(np.abs(Tm[big] / T[big] - 1) < 0.15)
    assert abs(np.mean(Tm[big] / T[big]) - 1) < 0.05


def test_replica_exchange_bookkeeping():
    off, x, t, T, lad = _agcu_small(ngeom=4, rungs=6)
    with _batch_agcu(hip.GH_MIXED, off, x, t) as b:
        b.set_temperatures(T)
        b.langevin(500, 0.002, 1.0, 3)
        x0, v0 = b.state()
        ladders = lad.copy()
        nacc = b.exchange(ladders, 0, 1234)
        x1, v1 = b.state()
    # rung order stays a permutation of each ladder's systems
    for l in range(lad.shape[0]):
        assert sorted(ladders[l]) == sorted(lad[l])
    assert np.array_equal(x0, x1)  # exchanges move temperatures, not atoms
    swapped = np.where(ladders != lad)
    assert nacc == len(swapped[0]) // 2
    # a swapped system's velocities scale by sqrt(T_new / T_old)
    for l, r in zip(*swapped):
        s = lad[l, r]
        new_rung = list(ladders[l]).index(s)
        f = np.sqrt(T[lad[l, new_rung]] / T[s])
        a0, a1 = off[s], off[s + 1]
        assert np.allclose(v1[a0:a1], v0[a0:a1] * f, rtol=1e-12, atol=0)


def _oracle_fire_all(off, x, s, workers=16):
    """The oracle's FP64 FIRE of every system, on host threads (the C oracle
    runs without the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(k):
        a0, a1 = off[k], off[k + 1]
        n = a1 - a0
        return O.fire(x[a0:a1], np.zeros(n), s, np.full(n, CL.MASS_AU))

    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, range(len(off) - 1)))


@pytest.mark.parametrize("precision", [hip.GH_MIXED, hip.GH_FP64], ids=["mixed", "fp64"])
def test_fire_relaxations_match_oracle(precision):
    """64 Au clusters (13..309 atoms, fcc spheres and Mackay fragments, 0.1 A
    perturbation): the batched FIRE (FP32 pair terms to 1e-2 eV/A, then FP64
    to 1e-4 eV/A; or FP64 throughout) against the oracle's FP64 FIRE from the
    same start.  Bars: every relaxation converges to a minimum (the oracle's
    FP64 forces there below 1.01e-4 eV/A, the reported energy the oracle's to
    1e-10); at least 95% reach the oracle's minimum, and within that basin the
    final energies agree to 1e-8 eV/atom."""
    off, x = CL.au_fire_batch(nclusters=64, seed=8)
    with hip.Batch(0, precision) as b:
        b.set_systems(off, x)
        b.set_smatb(1, CL.AU["A"], CL.AU["xi"], CL.AU["p"], CL.AU["q"], CL.AU["r0"], 4.078,
                    4.078 * np.sqrt(1.5))
        b.set_masses([CL.MASS_AU])
        r = b.fire(max_iter=20000, dt_max=0.005, alpha0=0.1, fswitch=1e-2, ftol=1e-4)
        xf, _ = b.state()
    assert r["converged"].all()
    s = O.smatb_params([(0, 0, CL.AU["A"], CL.AU["xi"], CL.AU["p"], CL.AU["q"], CL.AU["r0"],
                         4.078, 4.078 * np.sqrt(1.5))], 1)
    box = O.make_box(None)
    ro = _oracle_fire_all(off, x, s)
    n = np.diff(off)
    eo = np.array([q["epot"] for q in ro])
    assert all(q["converged"] for q in ro)
    for k in range(len(n)):
        a0, a1 = off[k], off[k + 1]
        # the GPU minimum is a minimum: FP64 oracle forces there are below tol
        nn, lst, _ = O.neighbours(xf[a0:a1], box, O.smatb_rcut(s) + 1.0)
        Fo, _, epo, _ = O.smatb_force(xf[a0:a1], np.zeros(n[k]), box, s, nn, lst)
        assert np.sqrt((Fo ** 2).sum(axis=1)).max() * cm.UI_EV < 1.01e-4
        assert r["epot"][k] == pytest.approx(epo * cm.UI_EV, rel=1e-10)
    de = np.abs(r["epot"] - eo) / n  # eV/atom
    same = de < 1e-6
    print("FIRE %s: %d/%d reach the oracle's minimum; within it max |dE| %.2e eV/atom; "
          "iterations mean %.0f (oracle %.0f)"
          % ("mixed" if precision == hip.GH_MIXED else "fp64", same.sum(), len(n),
             de[same].max(), np.mean(r["iters"]), np.mean([q["iters"] for q in ro])))
    assert same.mean() >= 0.95
    assert de[same].max() <= 1e-8


def _agcu_sizes(sizes, seed=21):
    """Ag-Cu fcc spheres of the given sizes (50:50 seeded labels, grouped by
    element), as the replica-exchange workload builds them."""
    rng = np.random.default_rng(seed)
    a = CL.AGCU["r0"] * np.sqrt(2.0)
    xs, ts, off = [], [], [0]
    for n in sizes:
        x = CL.fcc_sphere(int(n), a, rng.uniform(-0.5, 0.5, 3) * a)
        lab = np.zeros(int(n), np.int32)
        lab[rng.permutation(int(n))[: int(n) // 2]] = 1
        order = np.argsort(lab, kind="stable")
        xs.append(x[order])
        ts.append(lab[order])
        off.append(off[-1] + int(n))
    return np.array(off, np.int32), np.concatenate(xs), np.concatenate(ts)


# the replica-exchange workload's sizes reach 2000 atoms: systems above 1024
# atoms run 1024-thread workgroups on a second stream beside the rest
BIG = [1878, 38, 1100, 640, 1350, 95]


@pytest.mark.parametrize("prec,tol", [(hip.GH_FP64, 1e-10), (hip.GH_MIXED, 1e-5)],
                         ids=["fp64", "mixed"])
def test_batch_forces_with_systems_above_1024_atoms(prec, tol):
    off, x, t = _agcu_sizes(BIG)
    with _batch_agcu(prec, off, x, t) as b:
        F, ep = b.compute()
    s = _oracle_agcu()
    box = O.make_box(None)
    for k in range(len(off) - 1):
        a0, a1 = off[k], off[k + 1]
        nn, lst, mode = O.neighbours(x[a0:a1], box, O.smatb_rcut(s) + 1.0)
        Fo, eo, epo, _ = O.smatb_force(x[a0:a1], t[a0:a1], box, s, nn, lst)
        assert cm.force_err(F[a0:a1], Fo) <= tol, (k, a1 - a0)
        assert ep[k] == pytest.approx(epo * cm.UI_EV, rel=1e-10 if prec == 0 else 1e-7)


@pytest.mark.parametrize("n", [1100, 1878])
def test_batch_langevin_above_1024_atoms_equals_single_system_path(n):
    off, x, t = _agcu_sizes([n, 200])  # the big one on the second stream
    one = np.array([0, n], np.int32)
    with _batch_agcu(hip.GH_FP64, off, x, t) as b:
        b.set_temperatures([600.0, 300.0])
        ep, ek = b.langevin(100, 0.002, 1.0, 99)
        xb, vb = b.state()
    with hip.Context(0, hip.GH_FP64) as c:
        c.set_box(None)
        c.set_atoms(x[:n], t[:n])
        c.set_masses([CL.MASS_AG, CL.MASS_CU])
        c.set_smatb(2, *CL.agcu_params())
        c.integrator_ermak(0.002, 600.0, 1.0, rng=hip.GH_RNG_PHILOX, seed=99)
        c.run(100)
        xc, vc, _, _ = c.state()
        epc, ekc = c.energies()
    assert np.max(np.abs(xb[:n] - xc)) < 1e-9
    assert np.max(np.abs(vb[:n] - vc)) < 1e-7
    assert ep[0] == pytest.approx(epc, rel=1e-11)
    assert ek[0] == pytest.approx(ekc, rel=1e-9)


def test_replica_exchange_with_systems_above_1024_atoms():
    sizes = [1878, 1878, 1878, 1100, 1100, 1100, 60, 60, 60]  # three 3-rung ladders
    off, x, t = _agcu_sizes(sizes)
    T = np.array([300.0, 520.0, 900.0] * 3)
    lad = np.arange(9, dtype=np.int32).reshape(3, 3)
    with _batch_agcu(hip.GH_MIXED, off, x, t) as b:
        b.set_temperatures(T)
        b.langevin(300, 0.002, 1.0, 5)
        x0, v0 = b.state()
        ladders = lad.copy()
        nacc = b.exchange(ladders, 0, 77)
        x1, v1 = b.state()
    for l in range(3):
        assert sorted(ladders[l]) == sorted(lad[l])
    assert np.array_equal(x0, x1)
    swapped = np.where(ladders != lad)
    assert nacc == len(swapped[0]) // 2
    for l, r in zip(*swapped):
        s = lad[l, r]
        f = np.sqrt(T[lad[l, list(ladders[l]).index(s)]] / T[s])
        a0, a1 = off[s], off[s + 1]
        assert np.allclose(v1[a0:a1], v0[a0:a1] * f, rtol=1e-12, atol=0)


def test_batch_row_overflow_is_reported_and_resized():
    # a 300-atom Au cluster squeezed to 70 %: its rows outgrow the capacity the
    # batch sizes from close packing; the flags re-arm and the next call
    # re-sizes the rows and matches the oracle
    off, x0 = CL.au_fire_batch(nclusters=2, nmin=300, nmax=300, sigma=0.0, seed=3)
    x = x0 * 0.7
    p = [CL.AU["A"], CL.AU["xi"], CL.AU["p"], CL.AU["q"], CL.AU["r0"], 4.078,
         4.078 * np.sqrt(1.5)]
    with hip.Batch(0, hip.GH_FP64) as b:
        b.set_systems(off, x)
        b.set_smatb(1, *p)
        b.set_masses([CL.MASS_AU])
        with pytest.raises(hip.GhError) as e:
            b.compute()
        assert e.value.code == -4
        F, ep = b.compute()
    s = O.smatb_params([(0, 0) + tuple(p)], 1)
    box = O.make_box(None)
    for k in range(2):
        a0, a1 = off[k], off[k + 1]
        nn, lst, _ = O.neighbours(x[a0:a1], box, O.smatb_rcut(s) + 1.0)
        Fo, _, epo, _ = O.smatb_force(x[a0:a1], np.zeros(a1 - a0, np.int32), box, s, nn, lst)
        assert cm.force_err(F[a0:a1], Fo) <= 1e-10
