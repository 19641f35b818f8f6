"""GEMS itself on the device path: the drop-in at the reference's own plugin
boundary.

fortran/gems.mk builds GEMS from its sources with its force calls
(smatb_interact, src/TB.F90:457-586; lj/plj/slj/wca_interact,
src/Pairs.F90:207-560) handed to fortran/gems_plugin_hip.f90 and its
`dinamica` loop (dinamic, src/Programs.f90:114-167) to
fortran/gems_plugin_hip_md.f90 when GEMS_BACKEND=hip.  GEMS keeps its script
interpreter, its set-up commands (`fill`, `set tempgdist`, the lcg stream)
and its outputs; the steps run on the device between the steps GEMS writes
(the log says `# HIP dinamica N steps of M atoms on the device`).  With
GEMS_HIP_LOOP=0 GEMS integrates on the host and only the force calls go to
the device (gh_set_positions + gh_compute per step).

Workloads: the scripts that wrote the fixtures with the reference binary
(tests/golden/make_golden.py): md_script for au864_nve (864 Au, `prng lcg`
4242, 300 K, dt 2 fs, 200 steps), lj4000_nve (4000 LJ atoms, T* 0.72, dt
0.005 tau, 100 steps), lj32_ghost_nve (`box mic F`, 400 steps), au864_ermak
and agcu201_ermak (`evolve ermak 300|600 1`, 100 steps; the Ag-Cu cluster in
no box); nve_plj_script for examples/NVE/v_verlet.gms (100 He + 100 Ne from
`fill`, three plj groups, two of them self and one `with` group, 5000
steps) and the NVT example (the same gas under `evolve ermak 300 10`, 10^4
steps); examples/SMA-TB/bulk.gms and md.gms restated (GEMS builds the 4000-atom
Ag crystal itself; md.gms runs GEMS's `lbfgs` first); drift_main's script for
the 32,000-atom Au NVE trace (10^4 steps).

Bars: the energy columns within 1e-9 relative of the reference's 16-digit
columns (Ermak: of each column's largest entry) and final positions within
1e-8 A in FP64; mixed precision within 1e-6 relative and 1e-4 A (LJ liquid:
1e-5 and 1e-3 A, its trajectory is more chaotic).  The NVE example and the
32k trace keep the bars of their direct-context tests (test_gpu_plj,
test_gpu_drift).
"""
import os
import shutil
import subprocess
import tempfile
import time

import numpy as np
import pytest

from tests import common as cm
from tests.golden import make_golden as mg

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEMS_HIP = os.path.join(REPO, "fortran", "_build", "gems")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(GEMS_HIP),
                                 reason="fortran/_build/gems not built (needs the GEMS sources)")]

AU_A = 4.078

CASES = [("au864_nve", "fp64", 1e-9, 1e-8), ("au864_nve", "mixed", 1e-6, 1e-4),
         ("lj4000_nve", "fp64", 1e-9, 1e-8), ("lj4000_nve", "mixed", 1e-5, 1e-3),
         ("lj32_ghost_nve", "fp64", 1e-9, 1e-8), ("lj32_ghost_nve", "mixed", 1e-5, 1e-3),
         ("au864_ermak", "fp64", 1e-9, 1e-8), ("au864_ermak", "mixed", 1e-6, 1e-4),
         ("agcu201_ermak", "fp64", 1e-9, 1e-8), ("agcu201_ermak", "mixed", 1e-6, 1e-4)]


def _script(name, f, wd):
    au_rows = [("Au", "Au", mg.AU, AU_A, AU_A * np.sqrt(1.5))]
    if name == "au864_nve":
        return mg.md_script(wd, ["Au"] * len(f["x"]), f["x"], f["L"], (1, 1, 1), au_rows,
                            int(f["seed"]), "> sys\nset tempgdist 300.\n", "evolve v_verlet\n",
                            int(f["nsteps"]), int(f["each"]), dt=0.002)
    if name == "au864_ermak":
        return mg.md_script(wd, ["Au"] * len(f["x"]), f["x"], f["L"], (1, 1, 1), au_rows,
                            int(f["seed"]), "", "evolve ermak 300. 1.\n",
                            int(f["nsteps"]), int(f["each"]), dt=0.002)
    if name == "agcu201_ermak":
        r2, r3 = np.sqrt(2.0), np.sqrt(3.0)
        rows = [("Ag", "Ag", mg.AG, mg.AG["r0"] * r2, mg.AG["r0"] * r3),
                ("Cu", "Cu", mg.CU, mg.CU["r0"] * r2, mg.CU["r0"] * r3),
                ("Ag", "Cu", mg.AGCU, mg.AGCU["r0"] * r2, mg.AGCU["r0"] * r3)]
        first = cm.meta()["agcu_first_symbol"]
        other = "Cu" if first == "Ag" else "Ag"
        syms = [first if t == 0 else other for t in f["types"]]
        return mg.md_script(wd, syms, f["x"], None, (0, 0, 0), rows, int(f["seed"]), "",
                            "evolve ermak 600. 1.\n", int(f["nsteps"]), int(f["each"]),
                            dt=0.002)
    T = 0.72 / mg.KB_EV
    return mg.md_script(wd, ["Ar"] * len(f["x"]), f["x"], f["L"], (1, 1, 1),
                        "nb_dcut = 0.3\ninteract 1 under lj 1. 1. 2.5\n", int(f["seed"]),
                        "> sys\nset tempgdist %s\n" % mg.fmt(T), "evolve v_verlet\n",
                        int(f["nsteps"]), int(f["each"]), mass=1.0,
                        dt=0.005 * np.sqrt(1.0 / mg.EV_UI),
                        extra="box mic F\n" if "ghost" in name else "")


def _run(wd, script, precision, loop=True, timeout=240):
    """Runs the GEMS binary on script in wd; returns its log and wall time."""
    with open(os.path.join(wd, "in.gms"), "w") as fh:
        fh.write(script)
    env = dict(os.environ, GEMS_BACKEND="hip", GEMS_HIP_PRECISION=precision,
               GEMS_HIP_LOOP="1" if loop else "0")
    t0 = time.perf_counter()
    r = subprocess.run([GEMS_HIP, "in.gms"], cwd=wd, capture_output=True, text=True,
                       timeout=timeout, env=env)
    wall = time.perf_counter() - t0
    log = open(os.path.join(wd, "in.log")).read()
    assert r.returncode == 0 and "#-ERR->" not in log, log[-2000:]
    on_device = "on the device" in log
    assert on_device == loop, [ln for ln in log.splitlines() if "HIP" in ln]
    return log, wall


def _check_trace(name, E, X, f, etol, xtol):
    ref = f["energy"]
    assert E.shape == ref.shape
    if "ermak" in name:
        rel = np.max(np.abs(E - ref) / np.abs(ref).max(axis=0))
    else:
        rel = np.max(np.abs(E - ref) / np.abs(ref))
    d = X - f["x_final"]
    if "L" in f:
        d -= f["L"] * np.round(d / f["L"])
    dx = np.abs(d).max()
    print("%s: energy columns max rel diff %.3e, final positions max diff %.3e A"
          % (name, rel, dx))
    assert rel <= etol
    assert dx <= xtol


@pytest.mark.parametrize("name,precision,etol,xtol", CASES)
def test_gems_script_on_the_device_loop(name, precision, etol, xtol):
    f = cm.load(name)
    wd = tempfile.mkdtemp(prefix="gems_hip_")
    try:
        print("gems_hip run:", name, precision, wd, flush=True)
        _run(wd, _script(name, f, wd), precision)
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)
        X = mg.read_frames(os.path.join(wd, "X.xyz"))[-1]
        _check_trace(name, E, X, f, etol, xtol)
    finally:
        shutil.rmtree(wd)


@pytest.mark.parametrize("name", ["au864_nve", "lj32_ghost_nve", "au864_ermak"])
def test_gems_force_hooks_with_the_host_loop(name):
    """GEMS_HIP_LOOP=0: GEMS integrates; each force call goes to the device."""
    f = cm.load(name)
    wd = tempfile.mkdtemp(prefix="gems_hip_")
    try:
        print("gems_hip host-loop run:", name, wd, flush=True)
        _run(wd, _script(name, f, wd), "fp64", loop=False)
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)
        X = mg.read_frames(os.path.join(wd, "X.xyz"))[-1]
        _check_trace(name, E, X, f, 1e-9, 1e-8)
    finally:
        shutil.rmtree(wd)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gems_nve_example_three_plj_groups(precision):
    """examples/NVE/v_verlet.gms through GEMS: `fill`, `set tempgdist` on the
    host, the 5000 steps on the device with the three plj groups (He-He,
    Ne-Ne self groups, the He-Ne `with` group) in one context of two types."""
    f = np.load(cm.GOLDEN + "/nve_plj.npz")
    kat = np.loadtxt(cm.GOLDEN + "/ref_Energy.v_verlet.dat")
    wd = tempfile.mkdtemp(prefix="gems_hip_")
    try:
        s, _, _ = mg.nve_plj_script(None, 5000, "outfile :e name E.dat\noutfile :e cols energy 1\n"
                                                "outfile :e each 100\n")
        print("gems_hip run: v_verlet.gms", precision, wd, flush=True)
        _run(wd, s, precision)
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)
    finally:
        shutil.rmtree(wd)
    ref = f["energy"]
    unit = 10.0 ** (np.floor(np.log10(np.abs(kat))) - 6)
    assert E.shape == ref.shape
    print("%s: epot to step 2000 max rel %.2e; ekin/total max rel %.2e; last-digit units to "
          "step 1800 max %.2f" % (precision,
                                  np.max(np.abs(E[:20, 0] - ref[:20, 0]) / np.abs(ref[:20, 0])),
                                  np.max(np.abs(E[:20, 1:] - ref[:20, 1:]) / np.abs(ref[:20, 1:])),
                                  np.max(np.abs(E[:18] - kat[:18]) / unit[:18])))
    if precision == "fp64":
        assert np.all(np.abs(E[:20, 0] - ref[:20, 0]) <= 1e-6 * np.abs(ref[:20, 0]))
        assert np.all(np.abs(E[:20, 1:] - ref[:20, 1:]) <= 1e-9 * np.abs(ref[:20, 1:]))
        assert np.all(np.abs(E[:18] - kat[:18]) <= 0.6 * unit[:18])
    else:
        assert np.all(np.abs(E[:18, 2] - kat[:18, 2]) <= 1e-6 * np.abs(kat[:18, 2]))
        assert np.all(np.abs(E[:18, 0] - kat[:18, 0]) <= 1e-5)
    assert abs(np.median(E[-10:, 2]) - kat[0, 2]) <= 2e-5 * abs(kat[0, 2])


def _sma_tb_example(name, wd):
    """The reference's examples/SMA-TB/{bulk,md}.gms restated: the same commands
    with the script's expressions evaluated into literals, the parameter file
    and Ag13 start (tests/golden/baletto_2003.prm, ag13_coords.xyz) beside
    them, the energy file renamed E.dat."""
    shutil.copy(os.path.join(cm.GOLDEN, "baletto_2003.prm"), wd)
    head = "dimension 3\nprng lcg\nprng seed 123456\n"
    out = "outfile :e name E.dat\noutfile :e cols energy 1\noutfile :e each %d\n"
    if name == "bulk":
        fact = 2.89 / np.sqrt(0.5)
        s = head + ">< atom 0 0 0\n+< atom 0.5 0.5 0\n+< atom 0 0.5 0.5\n+< atom 0.5 0 0.5\n"
        s += "+< reply 1 0 0 10\n+< reply 0 1 0 10\n+< reply 0 0 1 10\n"
        s += "sys add Ag\nbox size 12 12 12\n> sys\nset move 1 1 1\n"
        s += "set expand %s %s %s\nbox expand %s %s %s\n" % ((mg.fmt(fact),) * 6)
        s += "group 1 add\ninteract 1 under tb read baletto_2003.prm\n" + out % 20
        s += "evolve ermak 1500 1.\ndinamica 220\n"
        return s
    shutil.copy(os.path.join(cm.GOLDEN, "ag13_coords.xyz"), os.path.join(wd, "coords.xyz"))
    s = head + ">< read coords.xyz\nsys add Ag\n> element Ag\ngroup 1 add\ntime step 0.001d0\n"
    s += "interact 1 under tb read baletto_2003.prm\n" + out % 100
    s += "lbfgs\n> sys\nevolve ermak 300 1\ndinamica 1000\n"
    return s


@pytest.mark.parametrize("name", ["bulk", "md"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gems_sma_tb_examples_known_answer(name, precision):
    """examples/SMA-TB/bulk.gms (4000 Ag atoms built by GEMS's own `atom`,
    `reply`, `expand` commands, Ermak 1500 K, 220 steps) and md.gms (Ag13:
    GEMS's own `lbfgs` with device forces, then Ermak 300 K, 1000 steps)
    through GEMS's binary with the loop on the device, against the
    reference's ref/Energy.bulk.dat and Energy.md.dat (7 digits): FP64 every
    entry within half a unit of the 7th digit (measured: every digit equal);
    mixed bulk within 2e-6 relative (measured 9e-7); mixed md within 1e-3 of
    the thermal scale (the largest ekin; measured 2.9e-4): the 13 atoms under
    light friction (1/ps, 1 fs steps) amplify the mixed form's 1e-6 force
    differences from 2e-5 eV at step 100 to 1e-4 eV at step 1000, where the
    FP64 form keeps every printed digit."""
    ref = np.loadtxt(cm.GOLDEN + "/ref_Energy.%s.dat" % name)
    wd = tempfile.mkdtemp(prefix="gems_hip_")
    try:
        print("gems_hip run: examples/SMA-TB", name, precision, wd, flush=True)
        log, _ = _run(wd, _sma_tb_example(name, wd), precision)
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)
    finally:
        shutil.rmtree(wd)
    assert E.shape == ref.shape
    rel = np.max(np.abs(E - ref) / np.abs(ref))
    therm = np.max(np.abs(E - ref)) / np.abs(ref[:, 1]).max()
    print("%s %s: max rel diff to the 7-digit file %.2e; max |dE| / max ekin %.2e"
          % (name, precision, rel, therm))
    if precision == "fp64":
        assert rel <= 5.01e-7
    elif name == "bulk":
        assert rel <= 2e-6
    else:
        assert therm <= 1e-3


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gems_nvt_ermak_example(precision):
    """examples/NVT/ermak.gms through GEMS: the NVE example's 200-atom gas
    (three plj groups in one device context) under `evolve ermak 300 10` for
    10^4 steps, energy and temperature every 100 (100 device chunks, the lcg
    stream handed to the device and back at each), against the reference's
    ref/Energy.ermak.dat: FP64 every entry within 0.6 of the file's last
    digit, mixed every entry within 2e-6 of the thermal scale (as the direct
    test, test_gpu_plj::test_nvt_ermak_example_known_answer)."""
    kat = np.loadtxt(cm.GOLDEN + "/ref_Energy.ermak.dat")
    wd = tempfile.mkdtemp(prefix="gems_hip_")
    try:
        s, _, _ = mg.nve_plj_script(None, 10000, "outfile :e name E.dat\n"
                                                 "outfile :e cols energy 1 temp 1\n"
                                                 "outfile :e each 100\n")
        s = s.replace("evolve v_verlet\n", "evolve ermak 300 10\n")
        print("gems_hip run: examples/NVT/ermak.gms", precision, wd, flush=True)
        _run(wd, s, precision)
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 4)
    finally:
        shutil.rmtree(wd)
    assert E.shape == kat.shape
    u = cm.last_digit_units(E, kat)
    scale = np.array([np.abs(kat[:, 1]).max()] * 3 + [np.abs(kat[:, 3]).max()])
    dev = np.abs(E - kat).max(axis=0) / scale
    print("nvt ermak via GEMS %s: max %.2f units of the last digit; per column max |dE| / "
          "thermal scale %s" % (precision, u.max(), dev))
    if precision == "fp64":
        assert np.all(u <= 0.6)
    else:
        assert np.all(dev <= 2e-6)


def au32k_script(wd, nsteps=10000, each=100):
    """drift_main's script (tests/golden/make_golden.py): 32,000 Au atoms,
    `prng lcg` 4242, 300 K, velocity Verlet dt 2 fs, energy every `each`."""
    x, L = mg.au_bulk(20)
    mg.write_xyz(os.path.join(wd, "x.xyz"), ["Au"] * len(x), x)
    mg.write_prm(os.path.join(wd, "p.prm"), [("Au", "Au", mg.AU, AU_A, AU_A * np.sqrt(1.5))])
    s = mg.head(4242) + "out float 25.16\n>< read x.xyz\nsys add\n"
    s += "box size %s %s %s\n" % ((mg.fmt(L),) * 3)
    s += "> sys\nset pbc T T T\n> sys\ngroup 1 add\ninteract 1 under tb read p.prm\n"
    s += "time step 0.002\n> sys\nset tempgdist 300.\n"
    s += "outfile :e name E.dat\noutfile :e cols energy 1\noutfile :e each %d\n" % each
    s += "> sys\nevolve v_verlet\ndinamica %d\n" % nsteps
    return s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gems_au32k_nve_trace(precision):
    """The 32,000-atom Au NVE run through GEMS against the reference's own
    10^4-step trace: every row within 1e-9 eV/atom (FP64) / 1e-6 (mixed), the
    drift within 1e-5 eV/atom."""
    ref = cm.load("au32k_nve_drift")
    n = 32000
    wd = tempfile.mkdtemp(prefix="gems_hip_")
    try:
        print("gems_hip run: au32k", precision, wd, flush=True)
        log, wall = _run(wd, au32k_script(wd, int(ref["nsteps"]), int(ref["each"])), precision)
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)
    finally:
        shutil.rmtree(wd)
    en = E
    assert en.shape == ref["energy"].shape
    tol = 1e-9 if precision == "fp64" else 1e-6
    dev = np.max(np.abs(en - ref["energy"])) / n
    drift = (en[-1, 2] - en[0, 2]) / n
    print("au32k via GEMS %s: max |dE|/N %.3e eV/atom, drift %.3e eV/atom, whole run %.1f s "
          "(%.3e atom-steps/s including GEMS's set-up)" % (precision, dev, drift, wall,
                                                          n * 1e4 / wall))
    assert dev <= tol
    assert abs(drift) <= 1e-5
