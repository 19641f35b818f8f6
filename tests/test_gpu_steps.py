"""The MD step loop's list upkeep and its stops (through the C ABI).

  * On the tiled velocity Verlet path the rebuild chain is enqueued only
    where a rebuild is expected; a due rebuild without its chain stops the
    run at that step and the host re-runs the step with the chain.  The
    trajectory is the one of the loop that enqueues the chain every step
    (per-step timing on): energy rows, rebuild count and final state equal
    bit for bit.
  * A neighbour row or a brick's halo that outgrows its capacity stops the
    call at that step with GH_EOVERFLOW and the step in the message
    (Neighbour.F90:71,1073-1074 leaves the reference's `mnb` overflow
    unchecked); the next call re-sizes and its forces match the oracle to
    1e-10 (FP64); a run continued after the stop completes that step first.
  * An atom outside a non-periodic box's grid is GH_ERANGE; moved back, the
    next call succeeds (the flag does not stay set).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import common as cm
from tests.test_gpu_parity import _oracle_forces, _setup

pytestmark = pytest.mark.gpu

hip = pytest.importorskip("gems_amd.hip")


def _md(name, prec, timing, T, dt, nsteps, each):
    f, c, s, box, skin = _setup(name, precision=prec)
    c.prng_lcg(777)
    c.set_gdist(T)
    c.integrator_vv(dt)
    c.set_timing(timing)
    en = c.run(nsteps, each=each)
    x, v, F, _ = c.state()
    nupd = c.stats()["nupd"]
    c.close()
    return en, x, v, F, nupd


@pytest.mark.parametrize("name,prec,T,dt", [
    ("au864_pbc", hip.GH_FP64, 1500.0, 0.002),
    ("au864_pbc", hip.GH_MIXED, 1500.0, 0.002),
    ("lj4000_pbc", hip.GH_FP64, 8355.203, 5.090233e-5),
], ids=["au864-fp64", "au864-mixed", "lj4000-fp64"])
def test_lazy_rebuild_chain_matches_every_step(name, prec, T, dt):
    every = _md(name, prec, True, T, dt, 400, 25)
    lazy = _md(name, prec, False, T, dt, 400, 25)
    print("%s rebuilds %d / %d" % (name, every[4], lazy[4]))
    assert every[4] >= 2  # the window holds rebuilds
    assert lazy[4] == every[4]
    for a, b in zip(every[:4], lazy[:4]):
        assert np.array_equal(a, b)


def _compressed(x, L, centre, radius, factor):
    d = x - centre
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(axis=1))
    sel = r < radius
    y = x.copy()
    y[sel] = centre + d[sel] * factor
    return y, int(sel.sum())


def test_row_overflow_in_compute_is_reported_and_resized():
    f, c, s, box, skin = _setup("au864_pbc")
    c.compute()
    cap0 = c.stats()["cap"]
    L = np.asarray(f["L"], float)
    # a 7 A sphere squeezed to 60 %: its atoms get ~4.6x the neighbours
    x, nsel = _compressed(f["x"], L, L / 2, 7.0, 0.6)
    assert nsel > 50
    c.set_positions(x)
    with pytest.raises(hip.GhError) as e:
        c.compute()
    assert e.value.code == -4 and "at step" in str(e.value)
    F, _, et = c.compute()  # re-sized
    assert c.stats()["cap"] > cap0
    g = dict(f)
    g["x"] = x
    Fo, _, epo, _, _ = _oracle_forces(g, s, box, skin)
    assert cm.force_err(F, Fo) <= 1e-10
    assert et == pytest.approx(epo, rel=1e-12)
    c.close()


def _implosion(nx=10, rho=0.5, v0=20.0):
    """LJ atoms on an fcc lattice at reduced density rho with a velocity field
    converging on the box centre (v = -v0 (x - c) / (L/2)): the centre's
    density rises until the rows outgrow their capacity."""
    a = (4.0 / rho) ** (1.0 / 3.0)
    x = cm.fcc(nx, a) + 0.25 * a
    L = nx * a
    v = -v0 * (x - L / 2) / (L / 2)
    return x, v, L


def test_overflow_mid_run_stops_at_the_step():
    x, v, L = _implosion()
    c = hip.Context(0, hip.GH_FP64)
    c.set_box([L] * 3, (1, 1, 1))
    c.set_atoms(x)
    c.set_masses([1.0])
    c.set_lj(1.0, 1.0, 2.5)
    c.set_skin(0.3)
    c.set_velocities(v)
    c.integrator_vv(5.090233e-5)
    # the compression goes on: each call may stop at a later overflow; the
    # state is then the start of the stopped step's force evaluation and the
    # next call completes that step first
    total, stops, caps = 3000, [], []
    while c.stats()["steps"] < total:
        try:
            c.run(total - c.stats()["steps"])
        except hip.GhError as e:
            msg = str(e)
            print(msg)
            assert e.code == -4 and "at step" in msg
            stop = int(msg.split("at step")[1].split(";")[0])
            assert c.stats()["steps"] == stop - 1
            assert not stops or stop > stops[-1]
            stops.append(stop)
            caps.append(c.stats()["cap"])
            assert len(stops) < 20
    assert stops  # the scenario overflows
    assert c.stats()["steps"] == total
    assert c.stats()["cap"] > caps[0]
    # forces at the final state against the oracle
    xs, vs, Fs, _ = c.state()
    assert np.all(np.isfinite(xs)) and np.all(np.isfinite(vs))
    F, _, et = c.compute()
    box = O.make_box([L] * 3, (1, 1, 1))
    nn, lst, _ = O.neighbours(xs, box, 2.5 + 0.3)
    Fo, _, epo = O.lj_force(xs, box, 1.0, 1.0, 2.5, nn, lst)
    assert cm.force_err(F, Fo) <= 1e-10
    c.close()


def test_atom_outside_grid_reported_then_recovers():
    f = cm.load("au864_pbc")
    L = np.asarray(f["L"], float)
    c = hip.Context(0, hip.GH_FP64)
    c.set_box(L, (0, 0, 0))
    c.set_atoms(f["x"])
    c.set_masses([cm.MASS_AU])
    c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
    c.compute()
    assert c.stats()["list_mode"] == 1
    x = f["x"].copy()
    x[5] = [-30.0, L[1] / 2, L[2] / 2]  # far outside the guard cells
    c.set_positions(x)
    with pytest.raises(hip.GhError) as e:
        c.compute()
    assert e.value.code == -5
    c.set_positions(f["x"])
    F, _, et = c.compute()
    s = O.smatb_params(cm.au_params(), 1)
    box = O.make_box(L, (0, 0, 0))
    nn, lst, _ = O.neighbours(f["x"], box, O.smatb_rcut(s) + 1.0)
    Fo, _, epo, _ = O.smatb_force(f["x"], f["types"], box, s, nn, lst)
    assert cm.force_err(F, Fo) <= 1e-10
    c.close()
