"""The cooperative list build (a wavefront per atom row, 64 candidates per
step; GEMS_HIP_LIST_BUILD=coop) against the per-thread build
(GEMS_HIP_LIST_BUILD=rows), each forced on one context: both write the
same rows in the same order, so forces, energies and whole trajectories
through rebuilds agree bit for bit in both precisions.  Row order decides
the FP64 summation order, so a reordered row shows up in the last bits of
the forces.

Cases: the golden Au 864 and LJ 4000 boxes (few bricks: several workgroups
per brick), Au 32,000 atoms at three brick shapes, an LJ fcc crystal with
4 x 4 x 2-cell bricks, whose candidate lists overflow one LDS batch (the
cells then run in two batches), an Au slab free in z (guard cells) and an Au
cluster in a large periodic box (empty cells and bricks).  The sets themselves are checked against
the oracle in test_gpu_parity.py, on the cooperative path.
"""
import os

import numpy as np
import pytest

from tests import common as cm

pytestmark = pytest.mark.gpu

hip = pytest.importorskip("gems_amd.hip")


def _make(x, types, L, pot, precision, rows, skin):
    os.environ["GEMS_HIP_LIST_BUILD"] = "rows" if rows else "coop"
    try:
        c = hip.Context(0, precision)
    finally:
        os.environ.pop("GEMS_HIP_LIST_BUILD", None)
    c.set_box(L, [1, 1, 1])
    c.set_atoms(x, types)
    if pot == "lj":
        c.set_masses([1.0])
        c.set_lj(1.0, 1.0, 2.5)
    else:
        c.set_masses([cm.MASS_AU])
        c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
    c.set_skin(skin)
    return c


def _pair(x, types, L, pot, precision, skin, tile=None):
    out = []
    for rows in (False, True):
        c = _make(x, types, L, pot, precision, rows, skin)
        if tile is not None:
            c.set_tile(tile)
        out.append(c)
    return out


def _same(a, b):
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def _check_built(cs):
    kinds = [c.stats()["list_build"] for c in cs]
    assert kinds == [2, 1], kinds  # cooperative, then per-thread


@pytest.mark.parametrize("name,pot,prec,T,dt", [
    ("au864_pbc", "au", hip.GH_FP64, 1500.0, 0.002),
    ("au864_pbc", "au", hip.GH_MIXED, 1500.0, 0.002),
    ("lj4000_pbc", "lj", hip.GH_FP64, 8355.203, 5.090233e-5),
    ("lj4000_pbc", "lj", hip.GH_MIXED, 8355.203, 5.090233e-5),
], ids=["au864-fp64", "au864-mixed", "lj4000-fp64", "lj4000-mixed"])
def test_cooperative_rows_match_through_rebuilds(name, pot, prec, T, dt):
    f = cm.load(name)
    skin = 0.3 if pot == "lj" else 1.0
    cs = _pair(f["x"], f["types"], f["L"], pot, prec, skin)
    res = []
    for c in cs:
        F, e, et = c.compute()
        c.prng_lcg(7)
        c.set_gdist(T)
        c.integrator_vv(dt)
        en = c.run(300, 25)
        res.append((F, e, np.array([et]), en) + tuple(c.state()))
        res[-1] += (np.array([c.stats()["nupd"]]),)
    _check_built(cs)
    print(name, "rebuilds", res[0][-1][0])
    assert res[0][-1][0] >= 3  # the window holds rebuilds
    _same(res[0], res[1])
    for c in cs:
        c.close()


def _au32k(seed=3):
    a = 4.078
    x = cm.fcc(20, a)
    x += np.random.default_rng(seed).normal(0.0, 0.05, x.shape)
    return x, np.zeros(len(x), np.int32), np.array([20 * a] * 3)


@pytest.mark.parametrize("tile", [(1, 1, 1), (4, 3, 3), (2, 2, 2)])
@pytest.mark.parametrize("prec", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_cooperative_rows_match_au32k_tiles(tile, prec):
    x, t, L = _au32k()
    cs = _pair(x, t, L, "au", prec, 1.0, tile=np.array(tile))
    res = [c.compute() for c in cs]
    _check_built(cs)
    assert cs[0].stats()["maxnn"] == cs[1].stats()["maxnn"]
    _same(res[0][:2], res[1][:2])
    assert res[0][2] == res[1][2]
    for c in cs:
        c.close()


def test_cooperative_rows_match_in_batches():
    # LJ fcc 24^3 (55,296 atoms, rho* 0.8442, 0.05 sigma jitter): 14 cells of
    # 2.88 sigma per axis, bricks of 4 x 4 x 2 cells hold ~650 atoms whose
    # candidate lists (~17,500 entries) exceed what the LDS keeps beside the
    # staged halo, so each brick runs its cells in two batches
    a = (4.0 / 0.8442) ** (1.0 / 3.0)
    x = cm.fcc(24, a)
    x += np.random.default_rng(5).normal(0.0, 0.05, x.shape)
    L = np.array([24 * a] * 3)
    for prec in (hip.GH_FP64, hip.GH_MIXED):
        cs = _pair(x, np.zeros(len(x), np.int32), L, "lj", prec, 0.3, tile=np.array([4, 4, 2]))
        res = [c.compute() for c in cs]
        _check_built(cs)
        _same(res[0][:2], res[1][:2])
        assert res[0][2] == res[1][2]
        for c in cs:
            c.close()


def _slab_and_cluster():
    a = 4.078
    x = cm.fcc(16, a, 16, 10)
    x += np.random.default_rng(11).normal(0.0, 0.05, x.shape)
    # a slab: periodic in x and y, free in z (guard cells below and above)
    slab = (x, np.array([16 * a, 16 * a, 10 * a + 12.0]), [1, 1, 0])
    # a cluster of the same lattice cut to a sphere in a large periodic box:
    # most cells and whole bricks are empty
    c0 = x.mean(axis=0)
    ball = x[np.sum((x - c0) ** 2, axis=1) < 22.0 ** 2] - c0 + 60.0
    cluster = (ball, np.array([120.0, 120.0, 120.0]), [1, 1, 1])
    return {"slab": slab, "cluster": cluster}


@pytest.mark.parametrize("case,tile", [("slab", None), ("cluster", None), ("cluster", (4, 4, 4))],
                         ids=["slab", "cluster", "cluster-4x4x4"])
def test_cooperative_rows_match_slab_and_sparse_box(case, tile):
    # 4 x 4 x 4-cell bricks at the cluster's surface are mostly empty, so all
    # 64 cells' candidate lists fit one batch (the batch bounds at their limit)
    x, L, pbc = _slab_and_cluster()[case]
    t = np.zeros(len(x), np.int32)
    for prec in (hip.GH_FP64, hip.GH_MIXED):
        cs = []
        for rows in (False, True):
            os.environ["GEMS_HIP_LIST_BUILD"] = "rows" if rows else "coop"
            try:
                c = hip.Context(0, prec)
            finally:
                os.environ.pop("GEMS_HIP_LIST_BUILD", None)
            c.set_box(L, pbc)
            c.set_atoms(x, t)
            c.set_masses([cm.MASS_AU])
            c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
            c.set_skin(1.0)
            if tile is not None:
                c.set_tile(np.array(tile))
            cs.append(c)
        res = [c.compute() for c in cs]
        assert cs[0].stats()["list_mode"] == 1  # the tiled cell path
        _check_built(cs)
        _same(res[0][:2], res[1][:2])
        assert res[0][2] == res[1][2]
        for c in cs:
            c.close()
