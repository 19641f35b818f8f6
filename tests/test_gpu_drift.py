"""NVE energy conservation against the reference's own 10^4-step trace.

Workload (tests/golden/make_golden.py, drift_main): 32,000 Au atoms (fcc
20^3 cells, a = 4.078 A, N(0, 0.02 A) jitter from seed 20260101, wrapped),
SMA-TB with rci = a and rce = a*sqrt(3/2), periodic, `prng lcg` seed 4242,
`set tempgdist 300`, velocity Verlet dt = 2 fs, skin 1 A, energy every 100
steps (write_energy columns: epot, ekin, etot in eV).  The reference rebuilt
its list at every step (nupd = 10001); this path rebuilds on the skin test.

Bars:
  * every row of the trace: FP64 within 1e-9 eV/atom of the reference's
    16-digit columns (measured: 6e-14), mixed within 1e-6 eV/atom (measured:
    9e-8); chaos would amplify last-bit differences over 20 ps, so the FP64
    bar leaves four decades of room;
  * drift |E(10^4) - E(100)| / N <= 1e-5 eV/atom for both precisions, and the
    largest excursion max |E(k) - E(100)| / N within three times the
    reference's own plus 1e-6 eV/atom.
"""
import numpy as np
import pytest

from tests import common as cm

pytestmark = pytest.mark.gpu

hip = pytest.importorskip("gems_amd.hip")


def _au32k():
    a = 4.078
    x = cm.fcc(20, a) + np.random.default_rng(20260101).normal(0.0, 0.02, (32000, 3))
    L = 20 * a
    return np.mod(x, L), L


def _trace(precision):
    ref = cm.load("au32k_nve_drift")
    x, L = _au32k()
    c = hip.Context(0, precision)
    c.set_box([L] * 3, (1, 1, 1))
    c.set_atoms(x)
    c.set_masses([cm.MASS_AU])
    c.set_smatb(1, *cm.matrices(cm.au_params(), 1))
    c.set_skin(1.0)
    c.prng_lcg(int(ref["seed"]))
    c.set_gdist(300.0)
    c.integrator_vv(0.002)
    en = c.run(int(ref["nsteps"]), each=int(ref["each"]))
    c.close()
    return en, ref["energy"], len(x)


def _drift(en, n):
    tot = en[:, 2]
    return (tot[-1] - tot[0]) / n, np.max(np.abs(tot - tot[0])) / n


@pytest.mark.parametrize("precision", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_nve_drift_against_the_reference_trace(precision):
    en, ref, n = _trace(precision)
    assert en.shape == ref.shape
    tol = 1e-9 if precision == hip.GH_FP64 else 1e-6  # eV/atom
    assert np.max(np.abs(en - ref)) / n <= tol
    d, dev = _drift(en, n)
    d_ref, dev_ref = _drift(ref, n)
    print("drift eV/atom: this %.3e (max excursion %.3e); reference %.3e (%.3e); "
          "max |etot - etot_ref| / N over the trace %.3e, max |epot - epot_ref| / N %.3e"
          % (d, dev, d_ref, dev_ref, np.max(np.abs(en[:, 2] - ref[:, 2])) / n,
             np.max(np.abs(en[:, 0] - ref[:, 0])) / n))
    assert abs(d) <= 1e-5
    assert dev <= 3.0 * dev_ref + 1e-6
