"""Seeded cluster geometries for the batched workloads.

* ``fcc_sphere(n, a, center)``: the n sites of an fcc lattice (constant a)
  nearest a centre point.
* ``mackay(n)``: the first n sites of a 309-atom Mackay icosahedron, ordered
  by distance from its centre (13, 55, 147 and 309 are closed shells).
* ``agcu_re_batch`` and ``au_fire_batch`` build the replica-exchange and FIRE
  batches used by bench.py and the tests.
"""
import numpy as np

MASS_AG = float(np.float32(107.87))  # default-real literals widened to FP64
MASS_CU = float(np.float32(63.55))
MASS_AU = float(np.float32(196.97))

AG = dict(A=0.1028, xi=1.178, p=10.928, q=3.139, r0=2.889)
CU = dict(A=0.0855, xi=1.224, p=10.960, q=2.278, r0=2.556)
AGCU = dict(A=0.0980, xi=1.2274, p=10.700, q=2.805, r0=2.7224)
AU = dict(A=0.2061, xi=1.790, p=10.229, q=4.036, r0=2.884)


def fcc_sphere(n, a, center):
    m = int(np.ceil((3.0 * n / (4.0 * np.pi * 4.0)) ** (1.0 / 3.0))) + 3
    r = np.arange(-m, m + 1, dtype=np.float64)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0, 0.5, 0.5], [0.5, 0, 0.5]])
    pts = (g[:, None, :] + base[None, :, :]).reshape(-1, 3) * a
    d = np.linalg.norm(pts - center, axis=1)
    idx = np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0], np.round(d, 9)))[:n]
    return pts[idx] - center


def _ico_vertices():
    phi = (1.0 + 5.0 ** 0.5) / 2.0
    v = []
    for s1 in (-1, 1):
        for s2 in (-1, 1):
            v += [(0, s1, s2 * phi), (s1, s2 * phi, 0), (s2 * phi, 0, s1)]
    v = np.array(v, float)
    return v / np.linalg.norm(v[0])


def mackay(n, d_nn=1.0):
    """First n sites (by distance, stable) of a 4-shell Mackay icosahedron
    with nearest-neighbour distance d_nn."""
    V = _ico_vertices()
    # edges and faces of the icosahedron
    dist = np.linalg.norm(V[:, None] - V[None], axis=-1)
    e = dist[dist > 1e-9].min()
    edges = [(i, j) for i in range(12) for j in range(i + 1, 12) if abs(dist[i, j] - e) < 1e-6]
    faces = [(i, j, k) for (i, j) in edges for k in range(j + 1, 12)
             if abs(dist[i, k] - e) < 1e-6 and abs(dist[j, k] - e) < 1e-6]
    pts = [np.zeros(3)]
    for k in range(1, 5):
        shell = set()
        P = [tuple(np.round(k * v, 9)) for v in V]
        for (i, j) in edges:
            for t in range(0, k + 1):
                shell.add(tuple(np.round(k * V[i] + t * (V[j] - V[i]) * k / k, 9)))
        for (i, j, l) in faces:
            for t1 in range(1, k):
                for t2 in range(1, k - t1):
                    p = k * V[i] + t1 * (V[j] - V[i]) + t2 * (V[l] - V[i])
                    shell.add(tuple(np.round(p, 9)))
        shell |= set(P)
        pts += [np.array(p) for p in sorted(shell)]
    pts = np.array(pts)
    # radial shell spacing: edge length of the unit icosahedron
    pts *= d_nn / e
    d = np.linalg.norm(pts, axis=1)
    idx = np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0], np.round(d, 9)))[:n]
    return pts[idx]


def agcu_re_batch(n_geom=64, rungs=8, nmin=38, nmax=2000, t_lo=300.0, t_hi=900.0, seed=4):
    """Replica-exchange batch: n_geom Ag-Cu clusters (log-uniform sizes, fcc
    spheres with a = r0(AgCu) sqrt2, random 50:50 labels) times `rungs`
    temperatures of the geometric ladder t_lo .. t_hi.  Returns
    (off, x, types, T, ladders) with system l*rungs + r = geometry l at rung r;
    types: 0 = Ag, 1 = Cu (atoms of each cluster grouped by element)."""
    rng = np.random.default_rng(seed)
    a = AGCU["r0"] * np.sqrt(2.0)
    sizes = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_geom))).astype(int)
    T = t_lo * (t_hi / t_lo) ** (np.arange(rungs) / max(rungs - 1, 1))
    xs, ts, off, temps, lad = [], [], [0], [], []
    for g, n in enumerate(sizes):
        c = rng.uniform(-0.5, 0.5, 3) * a
        x = fcc_sphere(int(n), a, c)
        lab = np.zeros(int(n), np.int32)
        lab[rng.permutation(int(n))[: int(n) // 2]] = 1
        order = np.argsort(lab, kind="stable")
        x, lab = x[order], lab[order]
        row = []
        for r in range(rungs):
            row.append(len(off) - 1)
            xs.append(x)
            ts.append(lab)
            off.append(off[-1] + int(n))
            temps.append(T[r])
        lad.append(row)
    return (np.array(off, np.int32), np.concatenate(xs), np.concatenate(ts),
            np.array(temps), np.array(lad, np.int32))


def agcu_params():
    """(nt x nt matrices) for internal types 0 = Ag, 1 = Cu; hetero rci/rce
    as r0 sqrt2 and r0 sqrt3 of the mixed r0."""
    rows = {(0, 0): AG, (1, 1): CU, (0, 1): AGCU, (1, 0): AGCU}
    out = [np.zeros((2, 2)) for _ in range(7)]
    for (i, j), P in rows.items():
        vals = [P["A"], P["xi"], P["p"], P["q"], P["r0"], P["r0"] * np.sqrt(2.0),
                P["r0"] * np.sqrt(3.0)]
        for k in range(7):
            out[k][i, j] = vals[k]
    return out


def au_fire_batch(nclusters=4096, nmin=13, nmax=309, sigma=0.1, seed=5):
    """FIRE batch: Au clusters with sizes uniform in [nmin, nmax], half fcc
    spheres (a = 4.078) and half Mackay-icosahedron fragments (d_nn = 2.884),
    perturbed by N(0, sigma) per coordinate.  Returns (off, x)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, nclusters)
    ico = mackay(309, 2.884)
    xs, off = [], [0]
    for k, n in enumerate(sizes):
        if k % 2 == 0:
            x = fcc_sphere(int(n), 4.078, rng.uniform(-0.5, 0.5, 3) * 4.078)
        else:
            x = ico[: int(n)].copy()
        x = x + rng.normal(0.0, sigma, x.shape)
        xs.append(x)
        off.append(off[-1] + int(n))
    return np.array(off, np.int32), np.concatenate(xs)
