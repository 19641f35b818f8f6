"""Writes the golden fixtures of this directory by running the reference GEMS
binary (oracle/_ref/gems, compiled from the reference's own sources by
oracle/ref.mk) on seeded inputs.  Run in the build container only:

    make -f oracle/ref.mk REF=<path to the GEMS sources>
    python tests/golden/make_golden.py
    python tests/golden/make_golden.py --plj <path to the GEMS sources>
    python tests/golden/make_golden.py --ghost
    python tests/golden/make_golden.py --nvt <path to the GEMS sources>

Inputs are written with literal 17-digit numbers (no `{}` expressions or
`$var$` expansion), so the reference reads exactly the values stored here.
Outputs: forces (`outfile fce`, kcal/mol/A, 13 digits), per-atom energies
(`outfile pes`, eV, 13 digits), energy columns (`out float 25.16`), the
neighbour-rebuild count from the log, and final positions (`outfile pos`).
Each fixture is an .npz with the inputs and the reference's outputs.
"""
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
GEMS = os.path.join(REPO, "oracle", "_ref", "gems")

# src/Constants.F90: ui_kcm = ui_ev * eV_kcm with eV_kcm a default-real literal
UI_EV = 1.0e2 * 1.0e2 * 1.6605402e-27 * (1.0 / 1.60219e-19)
EV_UI = 1.0 / UI_EV
UI_KCM = UI_EV * float(np.float32(23.06035))
KB_EV = 8.617385e-05

AU = dict(A=0.2061, xi=1.790, p=10.229, q=4.036, r0=2.884)
AG = dict(A=0.1028, xi=1.178, p=10.928, q=3.139, r0=2.889)
CU = dict(A=0.0855, xi=1.224, p=10.960, q=2.278, r0=2.556)
AGCU = dict(A=0.0980, xi=1.2274, p=10.700, q=2.805, r0=2.7224)


def fcc(nx, a):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0, 0.5, 0.5], [0.5, 0, 0.5]])
    g = np.array([[i, j, k] for k in range(nx) for j in range(nx) for i in range(nx)], float)
    return ((g[:, None, :] + base[None, :, :]).reshape(-1, 3)) * a


def fmt(v):
    return "%.17e" % v


def write_xyz(path, syms, x):
    with open(path, "w") as f:
        f.write("%d\n\n" % len(x))
        for s, p in zip(syms, x):
            f.write("%s %s %s %s\n" % (s, fmt(p[0]), fmt(p[1]), fmt(p[2])))


def write_prm(path, rows):
    with open(path, "w") as f:
        f.write("GUPTA\n")
        for (e1, e2, P, rci, rce) in rows:
            f.write("%s %s %s %s %s %s %s %s %s\n" % (
                e1, e2, fmt(P["A"]), fmt(P["xi"]), fmt(P["p"]), fmt(P["q"]), fmt(P["r0"]),
                fmt(rci), fmt(rce)))


def read_frames(path, ncol_skip=1):
    """Reads GEMS xyz-like outputs written by write_fce/write_pos."""
    frames = []
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0
    while k < len(lines):
        if not lines[k].strip():
            k += 1
            continue
        n = int(lines[k].split()[0])
        rows = []
        for ln in lines[k + 2:k + 2 + n]:
            rows.append([float(t) for t in ln.split()[ncol_skip:ncol_skip + 3]])
        frames.append(np.array(rows))
        k += 2 + n
    return frames


def read_pes(path, n):
    vals = np.loadtxt(path)
    return vals.reshape(-1, n, 4)


def run(workdir, script, timeout=3600):
    with open(os.path.join(workdir, "in.gms"), "w") as f:
        f.write(script)
    r = subprocess.run([GEMS, "in.gms"], cwd=workdir, capture_output=True, text=True,
                       timeout=timeout)
    log = open(os.path.join(workdir, "in.log")).read()
    if "#-ERR->" in log or r.returncode != 0:
        raise RuntimeError("reference run failed:\n" + log[-3000:])
    # "vecinos actualizados: N veces" (Main.F90:338)
    m = re.findall(r"vecinos actualizados:\s*(\d+)", log)
    return log, (int(m[-1]) if m else -1)


def head(seed=None):
    s = "dimension 3\n"
    if seed is not None:
        s += "prng lcg\nprng seed %d\n" % seed
    return s


def pot_lines(wd, pot):
    """pot: GEMS script lines (str) or SMA-TB rows written to a literal .prm."""
    if isinstance(pot, str):
        return pot
    write_prm(os.path.join(wd, "p.prm"), pot)
    return "interact 1 under tb read p.prm\n"


def forces_fixture(name, syms, x, L, pbc, pot_block, types, mass=None, extra=""):
    """Reference forces/energies of one configuration (`out state`)."""
    wd = tempfile.mkdtemp(prefix="gold_")
    try:
        write_xyz(os.path.join(wd, "x.xyz"), syms, x)
        pot_block = pot_lines(wd, pot_block)
        s = head()
        s += "out float 25.16\n"
        s += ">< read x.xyz\nsys add\n"
        if L is not None:
            s += "box size %s %s %s\n" % tuple(fmt(v) for v in L)
            s += "> sys\nset pbc %s %s %s\n" % tuple("T" if p else "F" for p in pbc)
        s += "> sys\n"
        if mass is not None:
            s += "set mass %s\n" % fmt(mass)
        s += "group 1 add\n" + extra + pot_block
        s += ("outfile :e name E.dat\noutfile :e cols energy 1\noutfile :e each 1\n"
              "outfile :f name F.xyz\noutfile :f fce 1\noutfile :f each 1\n"
              "outfile :p name P.dat\noutfile :p pes 1\noutfile :p each 1\n"
              "out state\n")
        log, _ = run(wd, s)
        F = read_frames(os.path.join(wd, "F.xyz"))[-1] / UI_KCM  # ui/A
        pes = read_pes(os.path.join(wd, "P.dat"), len(x))[-1]
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)[-1]
        out = dict(x=x, types=np.asarray(types, np.int32), F=F, e_atom_ev=pes[:, 3],
                   epot_ev=E[0], x_ref=pes[:, :3])
        if L is not None:
            out["L"] = np.asarray(L, float)
            out["pbc"] = np.asarray(pbc, np.int32)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, "n=%d epot=%.16e" % (len(x), E[0]), flush=True)
    finally:
        shutil.rmtree(wd)


def md_script(wd, syms, x, L, pbc, pot_block, seed, setup, evolve, nsteps, each, mass=None,
              dt=None, extra=""):
    """Writes the inputs of an MD fixture into wd and returns the GEMS script:
    `dinamica nsteps` with the energy columns every `each` steps (E.dat) and
    the final positions (X.xyz)."""
    write_xyz(os.path.join(wd, "x.xyz"), syms, x)
    pot_block = pot_lines(wd, pot_block)
    s = head(seed)
    s += "out float 25.16\n"
    s += ">< read x.xyz\nsys add\n"
    if L is not None:
        s += "box size %s %s %s\n" % tuple(fmt(v) for v in L)
        s += "> sys\nset pbc %s %s %s\n" % tuple("T" if p else "F" for p in pbc)
    s += "> sys\n"
    if mass is not None:
        s += "set mass %s\n" % fmt(mass)
    s += "group 1 add\n" + extra + pot_block
    if dt is not None:
        s += "time step %s\n" % fmt(dt)
    s += setup
    s += ("outfile :e name E.dat\noutfile :e cols energy 1\noutfile :e each %d\n"
          "outfile :x name X.xyz\noutfile :x pos 1\noutfile :x each %d\n" % (each, nsteps))
    s += "> sys\n" + evolve + "dinamica %d\n" % nsteps
    return s


def md_fixture(name, syms, x, L, pbc, pot_block, types, seed, setup, evolve, nsteps, each,
               mass=None, dt=None, extra=""):
    """Reference energy trace of `dinamica nsteps` (16 digits) + final state."""
    wd = tempfile.mkdtemp(prefix="gold_")
    try:
        s = md_script(wd, syms, x, L, pbc, pot_block, seed, setup, evolve, nsteps, each, mass, dt,
                      extra)
        log, nupd = run(wd, s)
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)
        X = read_frames(os.path.join(wd, "X.xyz"))[-1]
        out = dict(x=x, types=np.asarray(types, np.int32), energy=E, x_final=X, nupd=nupd,
                   seed=seed, nsteps=nsteps, each=each)
        if L is not None:
            out["L"] = np.asarray(L, float)
            out["pbc"] = np.asarray(pbc, np.int32)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, "n=%d lines=%d nupd=%d E0=%.16e" % (len(x), len(E), nupd, E[0, 2]),
              flush=True)
    finally:
        shutil.rmtree(wd)


def main():
    if not os.path.exists(GEMS):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    rng = np.random.default_rng(20260101)
    meta = {}

    # --- Au, Cleri-Rosato parameters, rci = a, rce = a*sqrt(1.5) -----------
    a = 4.078
    au_rows = [("Au", "Au", AU, a, a * np.sqrt(1.5))]
    meta["au_prm"] = dict(AU, rci=a, rce=a * np.sqrt(1.5))

    for nx, name in [(4, "au256_pbc"), (6, "au864_pbc")]:
        x = fcc(nx, a) + rng.normal(0.0, 0.02, size=(4 * nx ** 3, 3))
        L = [nx * a] * 3
        x = np.mod(x, L[0])
        forces_fixture(name, ["Au"] * len(x), x, L, (1, 1, 1), au_rows, [0] * len(x))

    # --- Ag-Cu cluster (free, no box) --------------------------------------
    ac = 0.5 * (2.889 + 2.556) * np.sqrt(2.0)
    g = fcc(8, ac)
    g -= g.mean(axis=0)
    d = np.linalg.norm(g, axis=1)
    sel = np.argsort(d, kind="stable")[:201]
    x = g[sel] + rng.normal(0.0, 0.05, size=(201, 3)) + 50.0
    types = rng.integers(0, 2, size=len(x))
    # Atoms are grouped by element (all Ag, then all Cu).  smatb_new assigns
    # ig%z(k)=n for every atom (TB.F90:197-202), so interleaved elements would
    # collapse onto one internal type in the reference; grouped, the map is
    # the first-appearance numbering.
    order = np.argsort(types, kind="stable")
    x, types = x[order], types[order]
    syms = ["Ag" if t == 0 else "Cu" for t in types]
    r2, r3 = np.sqrt(2.0), np.sqrt(3.0)
    agcu_rows = [("Ag", "Ag", AG, AG["r0"] * r2, AG["r0"] * r3),
                 ("Cu", "Cu", CU, CU["r0"] * r2, CU["r0"] * r3),
                 ("Ag", "Cu", AGCU, AGCU["r0"] * r2, AGCU["r0"] * r3)]
    meta["agcu_prm"] = [dict(P, e1=e1, e2=e2, rci=rci, rce=rce) for (e1, e2, P, rci, rce) in agcu_rows]
    # internal types follow first appearance in the group (TB.F90:194-215)
    first = types[0]
    itypes = np.where(types == first, 0, 1)
    meta["agcu_first_symbol"] = syms[0]
    forces_fixture("agcu201_free", syms, x, None, (0, 0, 0), agcu_rows, itypes)

    # --- LJ, 4000 atoms (fcc 10^3 at rho*=0.8442), rc = 2.5, skin 0.3 -------
    alj = (4.0 / 0.8442) ** (1.0 / 3.0)
    x = fcc(10, alj) + rng.normal(0.0, 0.02, size=(4000, 3))
    L = [10 * alj] * 3
    x = np.mod(x, L[0])
    lj_block = "nb_dcut = 0.3\ninteract 1 under lj 1. 1. 2.5\n"
    forces_fixture("lj4000_pbc", ["Ar"] * 4000, x, L, (1, 1, 1), lj_block, [0] * 4000, mass=1.0)
    meta["lj"] = dict(eps=1.0, sigma=1.0, rc=2.5, skin=0.3, a=alj)

    # --- trajectories ---------------------------------------------------------
    au = np.load(os.path.join(HERE, "au864_pbc.npz"))
    md_fixture("au864_nve", ["Au"] * len(au["x"]), au["x"], au["L"], (1, 1, 1),
               au_rows, au["types"], seed=4242,
               setup="> sys\nset tempgdist 300.\n", evolve="evolve v_verlet\n",
               nsteps=200, each=10, dt=0.002)
    md_fixture("au864_ermak", ["Au"] * len(au["x"]), au["x"], au["L"], (1, 1, 1),
               au_rows, au["types"], seed=99,
               setup="", evolve="evolve ermak 300. 1.\n", nsteps=100, each=10, dt=0.002)
    lj = np.load(os.path.join(HERE, "lj4000_pbc.npz"))
    tau = np.sqrt(1.0 / EV_UI)
    T_lj = 0.72 / KB_EV
    meta["lj_nve"] = dict(T=T_lj, dt=0.005 * tau)
    md_fixture("lj4000_nve", ["Ar"] * 4000, lj["x"], lj["L"], (1, 1, 1), lj_block, lj["types"],
               seed=777, setup="> sys\nset tempgdist %s\n" % fmt(T_lj),
               evolve="evolve v_verlet\n", nsteps=100, each=10, mass=1.0, dt=0.005 * tau)
    cl = np.load(os.path.join(HERE, "agcu201_free.npz"))
    md_fixture("agcu201_ermak", syms, cl["x"], None, (0, 0, 0), agcu_rows,
               cl["types"], seed=31337, setup="", evolve="evolve ermak 600. 1.\n",
               nsteps=100, each=10, dt=0.002)

    # --- the reference's own Ag13 KAT: lbfgs minimum as a 13-digit fixture ---
    ag13 = np.array([[float(t) for t in ln.split()[1:4]] for ln in
                     open(os.path.join(HERE, "ag13_coords.xyz")).read().splitlines()[2:15]])
    wd = tempfile.mkdtemp(prefix="gold_")
    try:
        write_xyz(os.path.join(wd, "x.xyz"), ["Ag"] * 13, ag13)
        shutil.copy(os.path.join(HERE, "baletto_2003.prm"), wd)
        s = ("dimension 3\nout float 25.16\n>< read x.xyz\nsys add Ag\n> element Ag\ngroup 1 add\n"
             "interact 1 under tb read baletto_2003.prm\n"
             "outfile :x name X.xyz\noutfile :x pos 1\noutfile :x each 1\n"
             "outfile :e name E.dat\noutfile :e cols energy 1\noutfile :e each 1\n"
             "lbfgs\nout state\n")
        run(wd, s)
        X = read_frames(os.path.join(wd, "X.xyz"))[-1]
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)[-1]
        np.savez_compressed(os.path.join(HERE, "ag13_lbfgs.npz"), x=X, epot_ev=E[0])
        print("ag13_lbfgs epot=%.16e" % E[0])
    finally:
        shutil.rmtree(wd)

    with open(os.path.join(HERE, "meta.json"), "w") as f:
        json.dump(meta, f, indent=1, default=float)


def au_bulk(nx, seed=20260101):
    """Au fcc nx^3 cells, a = 4.078 A, N(0, 0.02 A) jitter from `seed`, wrapped."""
    a = 4.078
    x = fcc(nx, a) + np.random.default_rng(seed).normal(0.0, 0.02, (4 * nx ** 3, 3))
    L = nx * a
    return np.mod(x, L), L


def drift_main(nsteps=10000, each=100):
    """NVE energy trace of the reference over `nsteps` at 32,000 Au atoms
    (20^3 cells, 300 K from lcg seed 4242, dt 2 fs, skin 1 A).  Stored without
    positions: au_bulk(20) regenerates them."""
    x, L = au_bulk(20)
    a = 4.078
    wd = tempfile.mkdtemp(prefix="gold_")
    try:
        write_xyz(os.path.join(wd, "x.xyz"), ["Au"] * len(x), x)
        write_prm(os.path.join(wd, "p.prm"), [("Au", "Au", AU, a, a * np.sqrt(1.5))])
        s = head(4242) + "out float 25.16\n>< read x.xyz\nsys add\n"
        s += "box size %s %s %s\n" % ((fmt(L),) * 3)
        s += "> sys\nset pbc T T T\n> sys\ngroup 1 add\ninteract 1 under tb read p.prm\n"
        s += "time step 0.002\n> sys\nset tempgdist 300.\n"
        s += "outfile :e name E.dat\noutfile :e cols energy 1\noutfile :e each %d\n" % each
        s += "> sys\nevolve v_verlet\ndinamica %d\n" % nsteps
        t0 = __import__("time").time()
        log, nupd = run(wd, s, timeout=4 * 3600)  # about 75 min on 6 cores
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)
        np.savez_compressed(os.path.join(HERE, "au32k_nve_drift.npz"), energy=E, nupd=nupd,
                            nsteps=nsteps, each=each, seed=4242, nx=20, jitter_seed=20260101)
        print("au32k_nve_drift lines=%d nupd=%d E0=%.16e E1=%.16e (%.0f s)"
              % (len(E), nupd, E[0, 2], E[-1, 2], __import__("time").time() - t0))
    finally:
        shutil.rmtree(wd)


def ghost_main():
    """A ghost-image periodic box (`box mic F`, Neighbour.F90:1520-1794) small
    enough that one atom sees several images of another: LJ 2^3 fcc cells (32
    atoms, L = 3.36 sigma < 2 rc = 5 sigma); the forces of one configuration
    and an NVE trace.  (SMA-TB has no such fixture: the reference sizes its
    band/eband arrays to the local atoms when the interaction is declared
    (TB.F90:240-241) and writes the ghosts' densities past them
    (TB.F90:444-448), which faults for boxes this small.)"""
    rng = np.random.default_rng(20261202)
    ghost = "box mic F\n"
    alj = (4.0 / 0.8442) ** (1.0 / 3.0)
    x = np.mod(fcc(2, alj) + rng.normal(0.0, 0.02, size=(32, 3)), 2 * alj)
    L = [2 * alj] * 3
    lj_block = "nb_dcut = 0.3\ninteract 1 under lj 1. 1. 2.5\n"
    forces_fixture("lj32_ghost", ["Ar"] * 32, x, L, (1, 1, 1), lj_block, [0] * 32, mass=1.0,
                   extra=ghost)
    md_fixture("lj32_ghost_nve", ["Ar"] * 32, x, L, (1, 1, 1), lj_block, [0] * 32, seed=777,
               setup="> sys\nset tempgdist %s\n" % fmt(0.72 / KB_EV), evolve="evolve v_verlet\n",
               nsteps=400, each=10, mass=1.0, dt=0.005 * np.sqrt(1.0 / EV_UI), extra=ghost)


def nve_plj_script(ref_dir, nsteps, outputs):
    """The reference's examples/NVE/v_verlet.gms (100 He + 100 Ne placed by
    `fill` from the lcg stream, shifted LJ for the three group pairs, VV) with
    the pair parameters as literals (the example's `{}` expressions evaluated
    here) and `outputs` in place of its energy file."""
    kc = 2.61144768e19 / 6.0221417930e23 * 1000.0  # $kcalmol_ev$ of the example
    e = {"He": 0.021270 * kc, "Ne": 0.086000 * kc, "HeNe": np.sqrt(0.086000 * 0.021270) * kc}
    sg = {"He": 1.4800, "Ne": 1.5300, "HeNe": 0.5 * (1.5300 + 1.4800)}
    s = ("dimension 3\nbox size 50 50 50\nprng lcg\nprng seed 123456\n"
         ">< fill 100 2.6\nsys add He\n>< fill 100 2.6\nsys add Ne\n"
         "> sys\nset pbc T T T\ngroup 1 add\n> element He\ngroup 2 add\n"
         "> element Ne\ngroup 3 add\nbox move\ntime step 0.002d0\n")
    s += "interact 2 with 2 plj %s %s\n" % (fmt(e["He"]), fmt(sg["He"]))
    s += "interact 3 with 3 plj %s %s\n" % (fmt(e["Ne"]), fmt(sg["Ne"]))
    s += "interact 2 with 3 plj %s %s\n" % (fmt(e["HeNe"]), fmt(sg["HeNe"]))
    s += "out float 25.16\n" + outputs
    s += "> sys\nset tempgdist 300\nevolve v_verlet\ndinamica %d\n" % nsteps
    return s, e, sg


def nve_plj_main(ref_dir):
    """nve_plj.npz: the state after step 1 of the NVE example (positions,
    velocities, 16 digits) and the 16-digit energy columns every 100 steps to
    step 5000 of the same run; ref_Energy.v_verlet.dat: the example's own
    known-answer file (7 digits), copied."""
    wd = tempfile.mkdtemp(prefix="gold_")
    try:
        outs = ("outfile :x name X.xyz\noutfile :x pos 1\noutfile :x each 1\n"
                "outfile :v name V.xyz\noutfile :v vel 1\noutfile :v each 1\n")
        sc, e, sg = nve_plj_script(ref_dir, 1, outs)
        run(wd, sc)
        syms = [ln.split()[0] for ln in open(os.path.join(wd, "X.xyz")).read().splitlines()[2:]
                if ln.strip()]
        x1 = read_frames(os.path.join(wd, "X.xyz"))[-1]
        v1 = read_frames(os.path.join(wd, "V.xyz"))[-1]
        sc, _, _ = nve_plj_script(ref_dir, 5000, "outfile :e name E.dat\noutfile :e cols energy 1\n"
                                                  "outfile :e each 100\n")
        log, nupd = run(wd, sc)
        E = np.loadtxt(os.path.join(wd, "E.dat")).reshape(-1, 3)
        kat = os.path.join(ref_dir, "examples", "NVE", "ref", "Energy.v_verlet.dat")
        shutil.copy(kat, os.path.join(HERE, "ref_Energy.v_verlet.dat"))
        K = np.loadtxt(kat)
        ulp = 10.0 ** (np.floor(np.log10(np.abs(K))) - 6)
        types = np.array([0 if t == "He" else 1 for t in syms], np.int32)
        eps = np.array([[e["He"], e["HeNe"]], [e["HeNe"], e["Ne"]]])
        sig = np.array([[sg["He"], sg["HeNe"]], [sg["HeNe"], sg["Ne"]]])
        np.savez_compressed(os.path.join(HERE, "nve_plj.npz"), x1=x1, v1=v1, types=types,
                            # Elements.F90:65,73: default-real (FP32) literals
                            mass=np.array([np.float32(4.0032), np.float32(20.183)], np.float64),
                            L=50.0, eps=eps, sigma=sig,
                            rc=10.0, skin=1.0, dt=0.002, energy=E, each=100, nsteps=5000,
                            nupd=nupd)
        print("nve_plj n=%d rows=%d nupd=%d; vs the example's own file: max %.2f units of its "
              "last digit" % (len(x1), len(E), nupd, np.max(np.abs(E - K) / ulp)))
    finally:
        shutil.rmtree(wd)


def nve_plj_checkpoints(ref_dir):
    """chp_plj_s100.chp and chp_plj_s200.chp: the reference's checkpoints
    (`checkpoint write`) of the NVE example at steps 100 and 200 of one run;
    xyz_plj_s100.xyz and energy_plj.dat: its position frame at step 100 and
    its energy columns (default format) every 50 steps, for the format
    readers and writers."""
    wd = tempfile.mkdtemp(prefix="gold_")
    try:
        outs = ("outfile :x name X.xyz\noutfile :x pos 1\noutfile :x each 100\n"
                "outfile :e name E.dat\noutfile :e cols energy 1\noutfile :e each 50\n")
        sc, _, _ = nve_plj_script(ref_dir, 100, outs)
        sc = sc.replace("out float 25.16\n", "")
        sc += "checkpoint write s100.chp\ndinamica 100\ncheckpoint write s200.chp\n"
        run(wd, sc)
        shutil.copy(os.path.join(wd, "s100.chp"), os.path.join(HERE, "chp_plj_s100.chp"))
        shutil.copy(os.path.join(wd, "s200.chp"), os.path.join(HERE, "chp_plj_s200.chp"))
        frames = open(os.path.join(wd, "X.xyz")).read().splitlines()
        with open(os.path.join(HERE, "xyz_plj_s100.xyz"), "w") as f:
            f.write("\n".join(frames[:202]) + "\n")
        shutil.copy(os.path.join(wd, "E.dat"), os.path.join(HERE, "energy_plj.dat"))
        print("nve_plj checkpoints written")
    finally:
        shutil.rmtree(wd)


def nvt_ermak_main(ref_dir):
    """chp_nvt_s0.chp: the reference's checkpoint of examples/NVT/ermak.gms
    (the NVE example's 100 He + 100 Ne, `evolve ermak 300 10`) written just
    before `dinamica`, i.e. after `fill` and `set tempgdist 300` have drawn
    from the lcg stream; ref_Energy.ermak.dat: the example's own energy file.
    The 16-digit rows of its first 200 steps are printed as a check that the
    literal-parameter script is the example's run."""
    wd = tempfile.mkdtemp(prefix="gold_")
    try:
        outs = "outfile :e name E.dat\noutfile :e cols energy 1\noutfile :e each 100\n"
        sc, _, _ = nve_plj_script(ref_dir, 200, outs)
        sc = sc.replace("evolve v_verlet\ndinamica 200\n",
                        "evolve ermak 300 10\ncheckpoint write s0.chp\ndinamica 200\n")
        run(wd, sc)
        shutil.copy(os.path.join(wd, "s0.chp"), os.path.join(HERE, "chp_nvt_s0.chp"))
        shutil.copy(os.path.join(ref_dir, "examples", "NVT", "ref", "Energy.ermak.dat"),
                    os.path.join(HERE, "ref_Energy.ermak.dat"))
        print(open(os.path.join(wd, "E.dat")).read())
    finally:
        shutil.rmtree(wd)


if __name__ == "__main__":
    if "--drift" in sys.argv:
        drift_main()
    elif "--nvt" in sys.argv:
        nvt_ermak_main(sys.argv[sys.argv.index("--nvt") + 1])
    elif "--ghost" in sys.argv:
        ghost_main()
    elif "--plj" in sys.argv:
        nve_plj_main(sys.argv[sys.argv.index("--plj") + 1])
        nve_plj_checkpoints(sys.argv[sys.argv.index("--plj") + 1])
    else:
        main()
