"""Several workgroups per brick in the force passes (each stages the brick's
halo and takes a contiguous share of its atoms; GEMS_HIP_CHUNKS=K forces K)
against one workgroup per brick: an atom's sums do not depend on which
workgroup runs it, so forces, energies and whole trajectories agree bit for
bit.

Cases: the golden LJ 4000 and Au 864 boxes through 200 steps with rebuilds in
both precisions (four lanes per atom; FP64 LJ is k_lj_tiled, mixed LJ k_lj_fx,
SMA-TB both passes), the two-type Ag-Cu solid solution (the multi-type
kernels), and 131,072-atom LJ and Au boxes (one lane per atom).
"""
import os

import numpy as np
import pytest

from tests import common as cm

pytestmark = pytest.mark.gpu

hip = pytest.importorskip("gems_amd.hip")


def _ctx(prec, k):
    os.environ["GEMS_HIP_CHUNKS"] = str(k)
    try:
        return hip.Context(0, prec)
    finally:
        os.environ.pop("GEMS_HIP_CHUNKS", None)


def _setup(c, x, types, L, pot, skin=None):
    c.set_box(L, [1, 1, 1])
    c.set_atoms(x, types)
    if pot == "lj":
        c.set_masses([1.0])
        c.set_lj(1.0, 1.0, 2.5)
        c.set_skin(0.3 if skin is None else skin)
    elif pot == "au":
        c.set_masses([cm.MASS_AU])
        c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
        c.set_skin(1.0 if skin is None else skin)
    else:
        rows = cm.agcu_params("Ag")
        c.set_masses(cm.agcu_masses("Ag"))
        c.set_smatb(2, *cm.matrices(rows, 2))
        c.set_skin(1.0 if skin is None else skin)


def _same(a, b):
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("name,pot,prec,T,dt", [
    ("lj4000_pbc", "lj", hip.GH_FP64, 8355.203, 5.090233e-5),
    ("lj4000_pbc", "lj", hip.GH_MIXED, 8355.203, 5.090233e-5),
    ("au864_pbc", "au", hip.GH_FP64, 1500.0, 0.002),
    ("au864_pbc", "au", hip.GH_MIXED, 1500.0, 0.002),
], ids=["lj4000-fp64", "lj4000-mixed", "au864-fp64", "au864-mixed"])
def test_workgroups_per_brick_match_through_rebuilds(name, pot, prec, T, dt):
    f = cm.load(name)
    res = {}
    for k in (1, 2, 3, 4):
        with _ctx(prec, k) as c:
            _setup(c, f["x"], f["types"], f["L"], pot)
            F, e, et = c.compute()
            assert c.stats()["chunks"] == k
            c.prng_lcg(7)
            c.set_gdist(T)
            c.integrator_vv(dt)
            en = c.run(200, 25)
            res[k] = (F, e, np.array([et]), en) + tuple(c.state())
            nupd = c.stats()["nupd"]
    assert nupd >= 2  # the window holds rebuilds
    for k in (2, 3, 4):
        _same(res[1], res[k])


def test_workgroups_per_brick_two_types():
    a = 3.87
    x = cm.fcc(10, a)
    rng = np.random.default_rng(21)
    x += rng.normal(0.0, 0.05, x.shape)
    t = (rng.random(len(x)) < 0.5).astype(np.int32)
    t[0] = 0
    L = np.array([10 * a] * 3)
    for prec in (hip.GH_FP64, hip.GH_MIXED):
        res = []
        for k in (1, 3):
            with _ctx(prec, k) as c:
                _setup(c, x, t, L, "agcu")
                res.append(c.compute())
        _same(res[0][:2], res[1][:2])
        assert res[0][2] == res[1][2]


@pytest.mark.parametrize("pot", ["lj", "au"])
def test_workgroups_per_brick_one_lane(pot):
    # 131,072 atoms (fcc 32^3): one lane per atom
    if pot == "lj":
        a = (4.0 / 0.8442) ** (1.0 / 3.0)
        jit = 0.05
    else:
        a = 4.078
        jit = 0.05
    x = cm.fcc(32, a)
    x += np.random.default_rng(23).normal(0.0, jit, x.shape)
    L = np.array([32 * a] * 3)
    t = np.zeros(len(x), np.int32)
    for prec in (hip.GH_FP64, hip.GH_MIXED):
        res = []
        for k in (1, 2):
            with _ctx(prec, k) as c:
                _setup(c, x, t, L, pot)
                res.append(c.compute())
                assert c.stats()["chunks"] == k
        _same(res[0][:2], res[1][:2])
        assert res[0][2] == res[1][2]


def test_autotuned_workgroups_per_brick():
    # the autotuner may split the LJ 4000 box's 45 bricks over 2 or 4
    # workgroups each (its 1M-atom boxes keep one); whatever it keeps, the
    # forces are those of its brick shape on one workgroup per brick, bit for
    # bit (mixed forces depend on the shape: fixed point about the brick centre)
    f = cm.load("lj4000_pbc")
    for prec in (hip.GH_FP64, hip.GH_MIXED):
        with hip.Context(0, prec) as a, _ctx(prec, 1) as b:
            _setup(a, f["x"], f["types"], f["L"], "lj")
            _setup(b, f["x"], f["types"], f["L"], "lj")
            shape, _ = a.autotune(reps=2)
            k = a.stats()["chunks"]
            assert k in (1, 2, 4)
            print("precision %d: %s cells, %d workgroups per brick" % (prec, shape, k))
            b.set_tile(shape)
            ra, rb = a.compute(), b.compute()
            assert b.stats()["chunks"] == 1
            _same(ra[:2], rb[:2])
            assert ra[2] == rb[2]
            a.set_tile(shape)  # a shape request resets the split
            a.compute()
            assert a.stats()["chunks"] == 1
