"""Shared helpers of the test-suite: golden fixtures and parameter tables."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

# src/Constants.F90 (evaluated in real(dp)); eV_kcm is a default-real literal
UI_EV = 1.0e2 * 1.0e2 * 1.6605402e-27 * (1.0 / 1.60219e-19)
EV_UI = 1.0 / UI_EV
KB_EV = 8.617385e-05
KB_UI = KB_EV * EV_UI
MASS_AU = float(np.float32(196.97))
MASS_AG = float(np.float32(107.87))
MASS_CU = float(np.float32(63.55))


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def meta():
    with open(os.path.join(GOLDEN, "meta.json")) as f:
        return json.load(f)


def au_params():
    m = meta()["au_prm"]
    return [(0, 0, m["A"], m["xi"], m["p"], m["q"], m["r0"], m["rci"], m["rce"])]


def agcu_params(first_symbol):
    """Rows with internal types numbered by first appearance (TB.F90:194-215)."""
    m = meta()
    tid = {"Ag": 0, "Cu": 1} if first_symbol == "Ag" else {"Cu": 0, "Ag": 1}
    rows = []
    for r in m["agcu_prm"]:
        rows.append((tid[r["e1"]], tid[r["e2"]], r["A"], r["xi"], r["p"], r["q"], r["r0"],
                     r["rci"], r["rce"]))
    return rows


def agcu_masses(first_symbol):
    return [MASS_AG, MASS_CU] if first_symbol == "Ag" else [MASS_CU, MASS_AG]


def matrices(rows, nt):
    """(i,j,A,xi,p,q,r0,rci,rce) rows -> symmetric nt x nt arrays."""
    out = [np.zeros((nt, nt)) for _ in range(7)]
    for r in rows:
        i, j = r[0], r[1]
        for k in range(7):
            out[k][i, j] = r[2 + k]
            out[k][j, i] = r[2 + k]
    return out


def fcc(nx, a, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0, 0.5, 0.5], [0.5, 0, 0.5]])
    g = np.array([[i, j, k] for k in range(nz) for j in range(ny) for i in range(nx)], float)
    return ((g[:, None, :] + base[None, :, :]).reshape(-1, 3)) * a


def force_err(F, Fref):
    """max |dF| / rms |F| (per component)."""
    rms = np.sqrt(np.mean(Fref ** 2))
    return np.max(np.abs(F - Fref)) / rms


# examples/LJ/{lj,slj}.gms: two H atoms in a 5.78 A non-periodic box, a scan
# of their distance with `out state` rows, then velocity Verlet from rest
LJ_EXAMPLE_BOX = 5.78
LJ_EXAMPLE_DT = 0.002 * 0.01018  # `time step {0.002*$tau_ps$}`, tau_ps 0.01018
MASS_H = float(np.float32(1.008))


def lj_example_scan(energy_at):
    """The rows of the scan (`out state` at distance 1.2, then 250 steps of
    +0.01 from 0.9) and the positions the dynamics starts from, replaying the
    script's `set move` additions in order."""
    x = np.array([[1.0, 1.0, 1.0], [2.2, 1.0, 1.0]])
    rows = [energy_at(x)]
    x[1, 0] += -1.2
    x[1, 0] += 0.9
    for _ in range(250):
        x[1, 0] += 0.01
        rows.append(energy_at(x))
    x[1, 0] += -2.5
    x[1, 0] += -0.9
    x[1, 0] += 1.2
    return np.array(rows), x


def last_digit_units(E, kat):
    """|E - kat| in units of the 7-significant-digit file's last digit."""
    unit = 10.0 ** (np.floor(np.log10(np.where(kat == 0, 1.0, np.abs(kat)))) - 6)
    unit = np.where(kat == 0, 1e-7, unit)
    return np.abs(E - kat) / unit


def ghost_supercell(x, L, reach):
    """Local atoms (wrapped into [0, L)) followed by every periodic image of
    them that lies within `reach` of the box: a free-boundary system whose
    first len(x) atoms feel every image within `reach`, as the reference's
    ghost atoms make them do (pbcfullghost, Neighbour.F90:1703-1752; the
    images there are first-shell only, which is every image within the cut
    plus skin when each side is at least that long).  For SMA-TB the reach is
    twice the cut, so the densities of the images within the cut are whole.
    Shifted by +reach + 1 so that all coordinates are positive.  Returns the
    coordinates and, per atom, the local atom it is an image of."""
    L = np.asarray(L, float)
    x = np.mod(np.asarray(x, float), L)
    k = int(np.ceil(reach / L.min())) + 1
    imgs, tags = [x], [np.arange(len(x))]
    for sz in range(-k, k + 1):
        for sy in range(-k, k + 1):
            for sx in range(-k, k + 1):
                if sx == sy == sz == 0:
                    continue
                y = x + np.array([sx, sy, sz]) * L
                keep = np.all((y >= -reach) & (y <= L + reach), axis=1)
                imgs.append(y[keep])
                tags.append(np.nonzero(keep)[0])
    return np.concatenate(imgs) + (reach + 1.0), np.concatenate(tags)


def sm1_example_scan(energy_at):
    """examples/LJ/sm1.gms: two H atoms start on one site; `out state`, then
    300 steps of +0.01 A with `out state`; then -2 A and atom 2 at 10 A/ps
    along x.  Returns the scan rows and the start of the dynamics."""
    x = np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 1.0]])
    rows = [energy_at(x)]
    for _ in range(300):
        x[1, 0] += 0.01
        rows.append(energy_at(x))
    x[1, 0] += -2.0
    v = np.zeros((2, 3))
    v[1, 0] = 10.0
    return np.array(rows), x, v


SM1_DT = 0.001  # `time step 0.001`
