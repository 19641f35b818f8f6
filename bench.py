"""Headline benchmark: SMA-TB (Gupta) bulk Au NVE at 1,000,188 atoms on one
MI355X, mixed precision, beside the plain FP64 GPU form and a CPU baseline.

A step is one velocity-Verlet timestep of the whole system: stepa (drift and
half kick), the skin test and any neighbour rebuild, the two SMA-TB passes
(density, then forces and energies) and stepb (half kick).

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  `value` = atoms * K / wall time of the K timed steps
(inputs already resident in HBM).  `roofline` is for the dominant kernel, the
mixed-precision force pass, timed with HIP events on the library's stream.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "SMA-TB atom-steps/s at 1M atoms (mixed vs FP64 GPU) + batched cluster-steps/s"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per atom (FP64 x 24 + type 4 + rho 8 in pass 1;
# x 24 + type 4 + rho/eband 8 + F 24 + e 8 in pass 2)
BYTES_PASS1 = 36
BYTES_PASS2 = 68
# the timed force pass also applies velocity Verlet to its atoms (fused): the
# update adds v read + write (48) and the next x (24) per atom, and a fused
# step stores no F or e (pass 2 then touches x 24 + type 4 + eband 8)
BYTES_FUSED_VV = 72
BYTES_PASS2_FUSED = 36

AU = dict(A=0.2061, xi=1.790, p=10.229, q=4.036, r0=2.884)
A_LAT = 4.078
MASS_AU = float(np.float32(196.97))


def fcc(nx, a):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0, 0.5, 0.5], [0.5, 0, 0.5]])
    i = np.arange(nx, dtype=np.float64)
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), axis=-1)[..., ::-1].reshape(-1, 3)
    return ((g[:, None, :] + base[None, :, :]).reshape(-1, 3)) * a


def workload(nx, seed=20260101):
    x = fcc(nx, A_LAT)
    x = x + np.random.default_rng(seed).normal(0.0, 0.02, x.shape)
    L = nx * A_LAT
    return np.mod(x, L), L


def make_ctx(hip, precision, x, L, skin, lcg_seed=4242, T=300.0, dt=0.002):
    c = hip.Context(0, precision)
    c.set_box([L] * 3, (1, 1, 1))
    c.set_atoms(x)
    c.set_masses([MASS_AU])
    c.set_smatb(1, AU["A"], AU["xi"], AU["p"], AU["q"], AU["r0"], A_LAT, A_LAT * np.sqrt(1.5))
    c.set_skin(skin)
    c.prng_lcg(lcg_seed)
    c.set_gdist(T)
    c.integrator_vv(dt)
    return c


def profiled(field, kernel="k_tb_force_fx", summary="k_tb_force_fx_au1m_counters.md"):
    """A per-launch figure of `kernel` from the committed rocprofv3 summary:
    traffic_bytes_per_launch (FETCH_SIZE and WRITE_SIZE, calibrated against a
    1 GiB device copy) or valu_insts_per_launch (SQ_INSTS_VALU)."""
    path = os.path.join(REPO, "profiles", summary)
    try:
        for ln in open(path):
            if ln.startswith("%s %s:" % (field, kernel)):
                return int(ln.split(":")[1]), os.path.relpath(path, REPO)
    except OSError:
        pass
    return None, None


def fixed_tile(c, spec):
    """Sets the brick shape 'bx,by,bz'; returns (shape, None) like autotune."""
    bd = tuple(int(v) for v in spec.split(","))
    c.set_tile(bd)
    return bd, None


def timed_run(c, steps, warmup, timing=False):
    """Wall time of `steps` MD steps after `warmup` untimed ones.  With
    timing, HIP events bracket every launch (gh_set_timing) and the per-pass
    durations come back as well; the events themselves add about 30 us per
    step, so the headline region runs without them."""
    c.run(warmup)
    st0 = c.stats()
    c.set_timing(timing)
    t0 = time.perf_counter()
    c.run(steps)  # returns after the stream is synchronised
    t1 = time.perf_counter()
    tm = c.timing()
    c.set_timing(False)
    st1 = c.stats()
    return t1 - t0, tm, st1["nupd"] - st0["nupd"], st1["launches"] - st0["launches"]


def cpu_baseline(x, L, skin, budget_s=20.0):
    """The CPU oracle (C restatement of the reference's OpenMP build) on the
    host cores, over as many steps of the same workload as fit the budget."""
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    s = O.smatb_params([(0, 0, AU["A"], AU["xi"], AU["p"], AU["q"], AU["r0"], A_LAT,
                         A_LAT * np.sqrt(1.5))], 1)
    box = O.make_box([L] * 3, (1, 1, 1))
    mass = np.full(len(x), MASS_AU)
    v = O.set_gdist(O.LCG(4242), 300.0, mass)
    md = O.MDRun(x, v, mass, box, pot="smatb", tb=s, types=np.zeros(len(x)), skin=skin,
                 integ="vv", dt=0.002, nthreads=cores, maxnb=96)
    md.interact()  # first force evaluation (not timed, as in dinamic)
    steps, t = 0, 0.0
    while t < budget_s and steps < 200:
        t0 = time.perf_counter()
        md.run(1, each=0)
        t += time.perf_counter() - t0
        steps += 1
    return dict(value=len(x) * steps / t, unit="atom-steps/s", cores=cores, kind="port",
                sample="%d VV steps of the same %d-atom Au workload (oracle/gems_oracle.c, "
                       "OpenMP, %d threads; reference semantics incl. its rebuild rule: "
                       "%d rebuilds)" % (steps, len(x), cores, md.nupd - 1))


def _pool_map(fn, items, cores):
    """fn over items on host threads (the C oracle runs without the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=cores) as ex:
        return list(ex.map(fn, items))


def _cores():
    cores = len(os.sched_getaffinity(0))
    return min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))


def cpu_md_baseline(x, L, pot, mass, dt, T, skin, budget_s, max_steps, what):
    """The oracle's velocity-Verlet loop (OpenMP on the host cores) over as
    many steps of a bulk workload as fit the budget."""
    from oracle import oracle as O
    cores = _cores()
    box = O.make_box([L] * 3, (1, 1, 1))
    m = np.full(len(x), mass)
    v = O.set_gdist(O.LCG(4242 if pot == "smatb" else 777), T, m)
    kw = dict(pot="smatb", tb=O.smatb_params([(0, 0, AU["A"], AU["xi"], AU["p"], AU["q"],
                                                AU["r0"], A_LAT, A_LAT * np.sqrt(1.5))], 1),
              maxnb=96) if pot == "smatb" else dict(pot="lj", lj=(1.0, 1.0, 2.5), maxnb=96)
    md = O.MDRun(x, v, m, box, types=np.zeros(len(x)), skin=skin, integ="vv", dt=dt,
                 nthreads=cores, **kw)
    md.interact()  # the first force evaluation (not timed, as in dinamic)
    steps, t = 0, 0.0
    while t < budget_s and steps < max_steps:
        t0 = time.perf_counter()
        md.run(10, each=0)
        t += time.perf_counter() - t0
        steps += 10
    return dict(value=len(x) * steps / t, unit="atom-steps/s", cores=cores, kind="port",
                sample="%d of the %s's VV steps (oracle/gems_oracle.c, OpenMP, %d threads, "
                       "the reference's rebuild rule: %d rebuilds)"
                       % (steps, what, cores, md.nupd - 1))


def cpu_re_baseline(budget_s):
    """The oracle's Ermak Langevin (ermak_a/b, Integration.F90:963-1031) of a
    stratified sample of the replica-exchange batch (one cluster per
    geometry, the rungs in turn), each cluster on one host thread."""
    from gems_amd import clusters as CL
    from oracle import oracle as O
    cores = _cores()
    off, x, t, T, lad = CL.agcu_re_batch()
    A, xi, p, q, r0, rci, rce = CL.agcu_params()
    s = O.smatb_params([(i, j, A[i, j], xi[i, j], p[i, j], q[i, j], r0[i, j], rci[i, j],
                         rce[i, j]) for i in range(2) for j in range(2)], 2)
    box = O.make_box(None)
    sample = [int(lad[g, g % lad.shape[1]]) for g in range(lad.shape[0])]
    runs = []
    for k in sample:
        a0, a1 = off[k], off[k + 1]
        m = np.where(t[a0:a1] == 0, CL.MASS_AG, CL.MASS_CU)
        md = O.MDRun(x[a0:a1], np.zeros((a1 - a0, 3)), m, box, pot="smatb", tb=s,
                     types=t[a0:a1], skin=1.0, integ="ermak", dt=0.002, T=float(T[k]),
                     gamma=1.0, rng=O.LCG(1000 + k), nthreads=1)
        md.interact()
        runs.append(md)
    chunk, steps, wall = 20, 0, 0.0
    while wall < budget_s and steps < 2000:
        t0 = time.perf_counter()
        _pool_map(lambda md: md.run(chunk, each=0), runs, cores)
        wall += time.perf_counter() - t0
        steps += chunk
    atoms = sum(md.n for md in runs)
    return dict(value=len(runs) * steps / wall, unit="cluster-steps/s", cores=cores, kind="port",
                atom_steps_per_s=atoms * steps / wall,
                sample="%d of the 512 replica clusters (one per geometry, %d atoms in all, "
                       "every rung represented), %d Ermak steps each, one host thread per "
                       "cluster over %d threads" % (len(runs), atoms, steps, cores))


def cpu_fire_sample(x, off, sample, cores):
    """The oracle's FP64 FIRE of the sampled clusters on host threads; returns
    (results, wall seconds)."""
    from gems_amd import clusters as CL
    from oracle import oracle as O
    s = O.smatb_params([(0, 0, AU["A"], AU["xi"], AU["p"], AU["q"], AU["r0"], A_LAT,
                         A_LAT * np.sqrt(1.5))], 1)

    def one(k):
        a0, a1 = off[k], off[k + 1]
        return O.fire(x[a0:a1], np.zeros(a1 - a0), s, np.full(a1 - a0, CL.MASS_AU),
                      max_iter=20000, dt_max=0.005, alpha0=0.1, ftol=1e-4)

    t0 = time.perf_counter()
    res = _pool_map(one, sample, cores)
    return res, time.perf_counter() - t0


def gems_script(nx, nsteps, each):
    """The Au NVE run as a GEMS script (fcc nx^3 cells, the workload's jitter,
    `prng lcg` 4242, `set tempgdist 300`, velocity Verlet dt 2 fs, skin 1 A,
    energy every `each` steps), with its input files; returns (files, script)."""
    x, L = workload(nx)
    xyz = "%d\n\n" % len(x) + "".join("Au %.17e %.17e %.17e\n" % tuple(r) for r in x)
    prm = "GUPTA\nAu Au " + " ".join("%.17e" % v for v in (
        AU["A"], AU["xi"], AU["p"], AU["q"], AU["r0"], A_LAT, A_LAT * np.sqrt(1.5))) + "\n"
    s = ("dimension 3\nprng lcg\nprng seed 4242\nout float 25.16\n>< read x.xyz\nsys add\n"
         "box size %.17e %.17e %.17e\n> sys\nset pbc T T T\n> sys\ngroup 1 add\n"
         "interact 1 under tb read p.prm\ntime step 0.002\n> sys\nset tempgdist 300.\n"
         "outfile :e name E.dat\noutfile :e cols energy 1\noutfile :e each %d\n"
         "> sys\nevolve v_verlet\ndinamica %d\n" % (L, L, L, each, nsteps))
    return {"x.xyz": xyz, "p.prm": prm}, s


def gems_driven(nx=20, nsteps=2000, each=100):
    """GEMS's own binary (fortran/_build/gems, GEMS_BACKEND=hip) on the Au 32k
    NVE script: GEMS reads the script, builds the system and draws the
    velocities on the host; its `dinamica` loop runs on the device in chunks
    between the energy rows.  Rate = atom-steps / wall time of the loop as
    GEMS's log reports it (state written back at every row)."""
    import re
    import subprocess
    import tempfile
    exe = os.path.join(REPO, "fortran", "_build", "gems")
    if not os.path.exists(exe):
        return None
    files, script = gems_script(nx, nsteps, each)
    out = {"workload": "GEMS script: Au SMA-TB fcc %d^3 cells = %d atoms, 300 K (lcg 4242), VV "
                       "dt 2 fs, skin 1 A, `dinamica %d`, energy row every %d steps"
                       % (nx, 4 * nx ** 3, nsteps, each)}
    for prec in ("fp64", "mixed"):
        with tempfile.TemporaryDirectory() as wd:
            for k, v in files.items():
                with open(os.path.join(wd, k), "w") as f:
                    f.write(v)
            with open(os.path.join(wd, "in.gms"), "w") as f:
                f.write(script)
            env = dict(os.environ, GEMS_BACKEND="hip", GEMS_HIP_PRECISION=prec)
            t0 = time.perf_counter()
            r = subprocess.run([exe, "in.gms"], cwd=wd, capture_output=True, text=True,
                               timeout=300, env=env)
            wall = time.perf_counter() - t0
            log = open(os.path.join(wd, "in.log")).read()
            m = re.search(r"dinamica done in ([0-9.]+) s \(\s*([0-9.Ee+-]+) atom-steps/s", log)
            if r.returncode != 0 or m is None:
                out[prec] = {"error": log[-500:]}
                continue
            out[prec] = {"atom_steps_per_s": float(m.group(2)), "loop_s": float(m.group(1)),
                         "process_s": wall,
                         "rows": len(np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3))}
            t = re.search(r"device steps ([0-9.]+) s, state copies ([0-9.]+) s, state into the "
                          r"atoms ([0-9.]+) s, GEMS output ([0-9.]+) s", log)
            if t:
                out[prec]["loop_parts_s"] = {k: float(v) for k, v in zip(
                    ("device_steps", "state_copies", "state_into_atoms", "gems_output"),
                    t.groups())}
    return out


def nve_drift(hip, precision, x, L, skin, nsteps, tile=None, each=100):
    """Total-energy drift of velocity-Verlet NVE over nsteps (eV/atom), on the
    energy columns of write_energy every `each` steps, measured as the
    reference's own trace is: from the first row (step `each`) to the last."""
    c = make_ctx(hip, precision, x, L, skin)
    if tile:
        c.set_tile(tile)
    t0 = time.perf_counter()
    en = c.run(nsteps, each=each)
    dt = time.perf_counter() - t0
    c.close()
    r = drift_stats(en, len(x))
    r.update(steps=nsteps, seconds=dt)
    return r


def drift_stats(en, n):
    tot = en[:, 2]
    return {"drift_ev_per_atom": float((tot[-1] - tot[0]) / n),
            "max_abs_dev_ev_per_atom": float(np.max(np.abs(tot - tot[0])) / n)}


def drift_vs_reference(hip):
    """The 32,000-atom NVE trace the reference itself wrote (10^4 steps, energy
    every 100; tests/golden/au32k_nve_drift.npz) next to this path's trace of
    the same run, FP64 and mixed."""
    path = os.path.join(REPO, "tests", "golden", "au32k_nve_drift.npz")
    if not os.path.exists(path):
        return None
    ref = np.load(path)
    nx = int(ref["nx"])
    x = fcc(nx, A_LAT) + np.random.default_rng(int(ref["jitter_seed"])).normal(
        0.0, 0.02, (4 * nx ** 3, 3))
    L = nx * A_LAT
    x = np.mod(x, L)
    out = {"workload": "Au 32,000 atoms (20^3 fcc cells, 0.02 A jitter), 300 K lcg seed 4242, "
                       "VV dt 2 fs, skin 1 A, 10^4 steps, energy every 100 steps",
           "reference": drift_stats(ref["energy"], len(x))}
    for name, prec in (("fp64", hip.GH_FP64), ("mixed", hip.GH_MIXED)):
        c = make_ctx(hip, prec, x, L, 1.0, lcg_seed=int(ref["seed"]))
        en = c.run(int(ref["nsteps"]), each=int(ref["each"]))
        c.close()
        out[name] = drift_stats(en, len(x))
        out[name]["max_abs_diff_vs_reference_ev_per_atom"] = float(
            np.max(np.abs(en[:, 2] - ref["energy"][:, 2])) / len(x))
    return out


def lj_bulk(hip, nx, steps, warmup=10, jitter=0.0, Tstar=0.72, tiles=None):
    """Lennard-Jones (lj_interact, Pairs.F90:207-279) on the LJ-liquid geometry
    of the LJ NVE configuration: fcc nx^3 cells at rho* = 0.8442 (4 nx^3
    atoms), rc = 2.5 sigma (unshifted), skin 0.3 sigma, eps = 1 eV, sigma =
    1 A, m = 1 amu, T* from the lcg stream (`set tempgdist`), velocity Verlet
    dt = 0.005 tau; mixed and FP64, each with its autotuned brick shape (or the
    shape tiles[name] names)."""
    a = (4.0 / 0.8442) ** (1.0 / 3.0)
    x = fcc(nx, a)
    if jitter > 0:
        x = x + np.random.default_rng(7).normal(0.0, jitter, x.shape)
    L = nx * a
    x = np.mod(x, L)
    tau = 0.01018047  # ps: sqrt(m sigma^2 / eps) in GEMS units
    T = Tstar / 8.617330337217213e-05  # K (kB in eV/K)
    out = {"workload": "LJ fcc %d^3 cells = %d atoms, rho* 0.8442, rc 2.5 sigma, skin 0.3 sigma, "
                       "T* %.2f, VV dt 0.005 tau, %d steps, periodic" % (nx, len(x), Tstar, steps)}
    F = {}
    for name, prec in (("mixed", hip.GH_MIXED), ("fp64", hip.GH_FP64)):
        c = hip.Context(0, prec)
        c.set_box([L] * 3, (1, 1, 1))
        c.set_atoms(x)
        c.set_masses([1.0])
        c.set_lj(1.0, 1.0, 2.5)
        c.set_skin(0.3)
        c.prng_lcg(4242)
        c.set_gdist(T)
        c.integrator_vv(0.005 * tau)
        if tiles and tiles.get(name):
            shape = tiles[name]
            c.set_tile(shape)
        else:
            shape, _ = c.autotune(reps=3)
        F[name] = c.compute()[0]
        dt, _, nupd, launches = timed_run(c, steps, warmup)
        _, tm, _, _ = timed_run(c, steps, 0, timing=True)
        wpb = c.stats()["chunks"]
        c.close()
        out[name] = {"atom_steps_per_s": len(x) * steps / dt, "ms_per_step": dt / steps * 1e3,
                     "rebuilds": nupd, "launches_per_step": launches / steps,
                     "force_pass_us": tm["force"][0] / max(tm["force"][1], 1) * 1e3,
                     "brick_cells": shape, "workgroups_per_brick": wpb}
    # the fused pass moves x 24 + v 48 + the next x 24 = 96 algorithmic B/atom
    t_us = out["mixed"]["force_pass_us"]
    ach = 96.0 * len(x) / (t_us * 1e-6) / 1e9
    traffic, src = profiled("traffic_bytes_per_launch", "k_lj_fx",
                            "k_lj_fx_lj1m_counters.md") if len(x) > 10 ** 6 else (None, None)
    out["roofline"] = {"kernel": "k_lj_fx (mixed, velocity Verlet fused)", "bound": "hbm",
                       "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": src,
                       "bytes_per_atom": 96, "avg_launch_us": t_us}
    if jitter > 0:  # on the perfect lattice the forces vanish and the ratio means nothing
        rms = float(np.sqrt(np.mean(F["fp64"] ** 2)))
        out["force_err_mixed_vs_fp64_rms_over_rms"] = float(
            np.sqrt(np.mean((F["mixed"] - F["fp64"]) ** 2)) / rms)
    return out


def au_32k(hip, steps=2000, warmup=20):
    """The Au 32k configuration: fcc 20^3 cells (32,000 atoms), 0.02 A jitter,
    300 K from the lcg stream, VV dt 2 fs, skin 1 A; mixed and FP64."""
    x, L = workload(20)
    out = {"workload": "Au SMA-TB fcc 20^3 cells = 32,000 atoms, 0.02 A jitter, 300 K, "
                       "VV dt 2 fs, skin 1 A, %d steps, periodic" % steps}
    for name, prec in (("mixed", hip.GH_MIXED), ("fp64", hip.GH_FP64)):
        c = make_ctx(hip, prec, x, L, 1.0)
        shape, _ = c.autotune(reps=3)
        dt, _, nupd, launches = timed_run(c, steps, warmup)
        wpb = c.stats()["chunks"]
        c.close()
        out[name] = {"atom_steps_per_s": len(x) * steps / dt, "ms_per_step": dt / steps * 1e3,
                     "rebuilds": nupd, "launches_per_step": launches / steps,
                     "brick_cells": shape, "workgroups_per_brick": wpb}
    return out


def batch_re(hip, rounds, swap_every=500):
    """Replica-exchange Langevin of 512 Ag-Cu clusters (64 geometries x 8
    temperatures, 300..900 K geometric), friction 1/ps, dt 2 fs, swaps every
    `swap_every` steps; one persistent launch per swap interval."""
    from gems_amd import clusters as CL
    off, x, t, T, lad = CL.agcu_re_batch()
    b = hip.Batch(0, hip.GH_MIXED)
    b.set_systems(off, x, t)
    b.set_smatb(2, *CL.agcu_params())
    b.set_masses([CL.MASS_AG, CL.MASS_CU])
    b.set_temperatures(T)
    b.langevin(swap_every, 0.002, 1.0, 1)  # warm-up interval (untimed)
    lad = lad.copy()
    acc = tries = 0
    t0 = time.perf_counter()
    for r in range(rounds):
        b.langevin(swap_every, 0.002, 1.0, 2 + r)
        acc += b.exchange(lad, r & 1, 1000 + r)
        tries += lad.shape[0] * ((lad.shape[1] - (r & 1)) // 2)
    dt = time.perf_counter() - t0
    b.close()
    nsys, ntot = len(off) - 1, int(off[-1])
    steps = rounds * swap_every
    return {"systems": nsys, "atoms": ntot, "steps": steps, "seconds": dt,
            "cluster_steps_per_s": nsys * steps / dt, "atom_steps_per_s": ntot * steps / dt,
            "swap_acceptance": acc / max(tries, 1),
            "workload": "512 Ag-Cu clusters (38..2000 atoms log-uniform, fcc spheres, 50:50), "
                        "8-rung ladder 300..900 K, Langevin (Ermak) gamma 1/ps dt 2 fs, "
                        "exchange every %d steps, mixed precision" % swap_every}


def batch_fire(hip, nclusters=4096, sample_every=64, cpu=True):
    """FIRE relaxation of 4096 Au clusters (13..309 atoms): FP32 pair terms
    until max|F| < 1e-2 eV/A, FP64 polish to 1e-4 eV/A.  The final energies
    of every `sample_every`-th cluster are compared with the oracle's FP64
    FIRE from the same start (timed on the host cores: the CPU baseline of
    this workload); the same batch relaxed in plain FP64 on the GPU is timed
    beside it."""
    from gems_amd import clusters as CL
    off, x = CL.au_fire_batch(nclusters)
    out = {}
    for prec in (hip.GH_MIXED, hip.GH_FP64):
        b = hip.Batch(0, prec)
        b.set_systems(off, x)
        b.set_smatb(1, AU["A"], AU["xi"], AU["p"], AU["q"], AU["r0"], A_LAT, A_LAT * np.sqrt(1.5))
        b.set_masses([MASS_AU])
        t0 = time.perf_counter()
        r = b.fire(max_iter=20000, dt_max=0.005, alpha0=0.1, fswitch=1e-2, ftol=1e-4)
        out[prec] = (time.perf_counter() - t0, r)
        b.close()
    dt, r = out[hip.GH_MIXED]
    dt64, r64 = out[hip.GH_FP64]
    n = np.diff(off)
    res = {"clusters": nclusters, "atoms": int(off[-1]), "seconds": dt,
           "clusters_relaxed_per_s": nclusters / dt,
           "converged": float(np.mean(r["converged"])),
           "iters_mean": float(np.mean(r["iters"])), "iters_max": int(np.max(r["iters"])),
           "fp64_gpu_seconds": dt64, "fp64_gpu_clusters_relaxed_per_s": nclusters / dt64,
           "workload": "4096 Au clusters, sizes uniform 13..309, half fcc spheres / half "
                       "Mackay fragments, 0.1 A Gaussian perturbation; FIRE dt_max 5 fs, "
                       "alpha0 0.1, FP32 pair terms to 1e-2 eV/A then FP64 to 1e-4 eV/A"}
    if cpu:
        sample = list(range(0, nclusters, sample_every))
        cores = _cores()
        ro, wall = cpu_fire_sample(x, off, sample, cores)
        eo = np.array([q["epot"] for q in ro])
        de = np.abs(r["epot"][sample] - eo) / n[sample]
        same = de < 1e-6
        res["vs_oracle_fp64_fire"] = {
            "clusters": len(sample), "same_minimum": float(np.mean(same)),
            "max_final_energy_err_ev_per_atom_same_minimum":
                float(de[same].max()) if same.any() else None,
            "note": "every %d-th cluster relaxed by the oracle's FP64 FIRE from the same "
                    "start; same minimum = |dE| < 1e-6 eV/atom" % sample_every}
        res["cpu_baseline"] = dict(
            value=len(sample) / wall, unit="clusters relaxed/s", cores=cores, kind="port",
            sample="the same %d clusters relaxed by oracle/gems_oracle.c's FP64 FIRE, one "
                   "host thread per cluster over %d threads" % (len(sample), cores))
    return res


def copy_probe(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size: its FETCH_SIZE/WRITE_SIZE
    counters calibrate the HBM traffic read for the force kernels."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=63)
    ap.add_argument("--skin", type=float, default=1.0)
    ap.add_argument("--tile", default=None,
                    help="brick shape 'bx,by,bz' for the mixed run instead of the autotuner "
                         "(profiling runs fix it so that every run launches the same grids)")
    ap.add_argument("--tile64", default=None, help="the same for the FP64 run")
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-gems", action="store_true")
    ap.add_argument("--drift-steps", type=int, default=10000)
    ap.add_argument("--re-rounds", type=int, default=4)
    ap.add_argument("--fire-clusters", type=int, default=4096)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates HBM counters)")
    ap.add_argument("--lj-tile", default=None,
                    help="with --lj-only: brick shapes 'bx,by,bz[:bx,by,bz]' of the 1M LJ leg "
                         "(mixed[:FP64]) instead of the autotuner")
    ap.add_argument("--lj-only", action="store_true",
                    help="only run the Lennard-Jones legs (profiling runs of k_lj_fx, k_lj_tiled)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.lj_only:
        from gems_amd import hip
        tiles = None
        if args.lj_tile:  # 'bx,by,bz:bx,by,bz' = the mixed and FP64 shapes of the 1M leg
            mx, _, f64 = args.lj_tile.partition(":")
            tiles = {"mixed": [int(v) for v in mx.split(",")],
                     "fp64": [int(v) for v in f64.split(",")] if f64 else None}
        print(json.dumps({"lj_4000": lj_bulk(hip, 10, 1000),
                          "lj_1m": lj_bulk(hip, 64, 100, jitter=0.01, Tstar=1.44, tiles=tiles)}))
        return
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported (the path's "
              "sharding over several GPUs is designed, not built)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from gems_amd import hip

    x, L = workload(args.nx)
    n = len(x)

    # mixed precision (the headline)
    c = make_ctx(hip, hip.GH_MIXED, x, L, args.skin)
    tile_mx = fixed_tile(c, args.tile) if args.tile else c.autotune(reps=3)
    F_mx, _, et_mx = c.compute()
    e0 = c.energies()
    dt_mx, _, nupd_mx, launches_mx = timed_run(c, args.steps, args.warmup)
    e1 = c.energies()
    dt_ev, tm_mx, _, _ = timed_run(c, args.steps, 0, timing=True)
    c.close()

    # plain FP64 GPU form, same input
    k64 = args.fp64_steps or max(10, args.steps // 2)
    c = make_ctx(hip, hip.GH_FP64, x, L, args.skin)
    tile_64 = fixed_tile(c, args.tile64) if args.tile64 else c.autotune(reps=3)
    F_64, _, et_64 = c.compute()
    dt_64, _, nupd_64, launches_64 = timed_run(c, k64, max(2, args.warmup // 2))
    _, tm_64, _, _ = timed_run(c, k64, 0, timing=True)
    c.close()

    value = n * args.steps / dt_mx
    value64 = n * k64 / dt_64
    rms = float(np.sqrt(np.mean(F_64 ** 2)))
    err_rms = float(np.sqrt(np.mean((F_mx - F_64) ** 2)) / rms)
    err_max = float(np.max(np.abs(F_mx - F_64)) / rms)

    ms_force, cnt_force = tm_mx["force"]
    ms_rho, cnt_rho = tm_mx["rho"]
    t_force = ms_force / cnt_force * 1e-3
    t_rho = ms_rho / cnt_rho * 1e-3
    bytes_force = BYTES_PASS2_FUSED + BYTES_FUSED_VV
    achieved = bytes_force * n / t_force / 1e9
    traffic, traffic_src = profiled("traffic_bytes_per_launch")
    valu_insts, _ = profiled("valu_insts_per_launch")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_mx / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair terms, f64 positions/accumulation",
        "data": "synthetic",
        "config": {
            "workload": "Au SMA-TB bulk NVE, %d atoms (%d^3 fcc cells, a=4.078 A, 0.02 A "
                        "seeded jitter, 300 K Maxwell-Boltzmann via the reference lcg "
                        "stream), velocity Verlet dt=2 fs, skin %.2f A, periodic"
                        % (n, args.nx, args.skin),
            "atoms": n,
            "precision": "mixed",
        },
        "roofline": {
            "kernel": "k_tb_force_fx (SMA-TB pass 2 + fused velocity Verlet, mixed, LDS tiles)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "bytes_per_atom": bytes_force,
            "avg_launch_us": t_force * 1e6,
            "valu_issue": None if valu_insts is None else {
                "insts_per_launch": valu_insts,
                "issue_us": valu_insts * 4 / (1024 * 2.4e3),
                "frac_of_launch": valu_insts * 4 / (1024 * 2.4e3) / (t_force * 1e6),
                "note": "SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x 2.4 GHz): the pass is bound "
                        "by FP32/FP64 vector issue of the pair terms, not by HBM"},
            "timed_with": "HIP events around every launch, over a second timed region of "
                          "the same %d steps (%.4f ms/step with the events)"
                          % (args.steps, dt_ev / args.steps * 1e3),
        },
        "pass1": {"kernel": "k_tb_rho_fx", "avg_launch_us": t_rho * 1e6,
                  "achieved_GBs": BYTES_PASS1 * n / t_rho / 1e9},
        "fp64": {"value": value64, "steps": k64, "ms_per_step": dt_64 / k64 * 1e3,
                 "force_pass_us": tm_64["force"][0] / tm_64["force"][1] * 1e3,
                 "rho_pass_us": tm_64["rho"][0] / tm_64["rho"][1] * 1e3,
                 "rebuilds": nupd_64, "launches_per_step": launches_64 / k64},
        "mixed_vs_fp64": value / value64,
        "mixed_vs_fp64_force": {
            "force_evaluation": (None if tile_mx[1] is None or tile_64[1] is None
                                 else tile_64[1] / tile_mx[1]),
            "passes_in_md_step": ((tm_64["rho"][0] / tm_64["rho"][1]
                                   + tm_64["force"][0] / tm_64["force"][1])
                                  / (ms_rho / cnt_rho + ms_force / cnt_force)),
            "note": "force_evaluation: one gh_compute (both SMA-TB passes) at each form's "
                    "tuned brick shape; passes_in_md_step: the two passes of an MD step "
                    "(velocity Verlet fused into the second), HIP-event timed; the whole "
                    "step (mixed_vs_fp64) also holds the list upkeep both forms share"},
        "force_err_vs_fp64": {"rms_over_rms": err_rms, "max_over_rms": err_max},
        "epot_rel_err_vs_fp64": abs(et_mx / et_64 - 1.0),
        "energy_change_ev_per_atom": (e1[0] + e1[1] - e0[0] - e0[1]) / n,
        "rebuilds": nupd_mx,
        "launches_per_step": launches_mx / args.steps,
        "rebuild_note": "the timed window holds %d neighbour rebuilds (one every %.0f steps); "
                        "the rebuild chain is enqueued only at steps where one is expected"
                        % (nupd_mx, args.steps / max(nupd_mx, 1)),
        "tiles": {"mixed": {"brick_cells": tile_mx[0], "force_eval_us": tile_mx[1]},
                  "fp64": {"brick_cells": tile_64[0], "force_eval_us": tile_64[1]}},
        "step_breakdown_us": {k: v[0] / max(v[1], 1) * 1e3 for k, v in tm_mx.items()},
    }
    if args.drift_steps > 0:
        out["nve_drift"] = {
            "mixed": nve_drift(hip, hip.GH_MIXED, x, L, args.skin, args.drift_steps, tile_mx[0]),
            "fp64": nve_drift(hip, hip.GH_FP64, x, L, args.skin, args.drift_steps, tile_64[0]),
            "workload": "the same 1,000,188-atom Au NVE run, energy every 100 steps",
            "vs_reference_32k": drift_vs_reference(hip),
        }
    if not args.no_gems:
        out["gems_driven"] = gems_driven()
    if not args.no_batch:
        out["lj_4000"] = lj_bulk(hip, 10, 1000)
        out["lj_1m"] = lj_bulk(hip, 64, 100, jitter=0.01, Tstar=1.44)
        out["au_32k"] = au_32k(hip)
        out["batched_re"] = batch_re(hip, args.re_rounds)
        out["batched_fire"] = batch_fire(hip, args.fire_clusters, cpu=not args.no_cpu)
        if not args.no_cpu:
            b = args.cpu_budget / 2
            alj = (4.0 / 0.8442) ** (1.0 / 3.0)
            xl = np.mod(fcc(10, alj), 10 * alj)
            out["lj_4000"]["cpu_baseline"] = cpu_md_baseline(
                xl, 10 * alj, "lj", 1.0, 0.005 * 0.01018047, 0.72 / 8.617330337217213e-05,
                0.3, b, 1000, "LJ 4000-atom workload")
            x32, L32 = workload(20)
            out["au_32k"]["cpu_baseline"] = cpu_md_baseline(
                x32, L32, "smatb", MASS_AU, 0.002, 300.0, 1.0, b, 2000,
                "Au 32,000-atom workload")
            out["batched_re"]["cpu_baseline"] = cpu_re_baseline(b)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x, L, args.skin, args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
