"""ctypes wrapper around the C restatement of the GEMS hot path (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/, the
smoke check and the CPU-baseline leg of bench.py import this module; the
product package never does.

Arrays follow the reference's per-atom layout: positions ``x`` are (n, 3)
float64 in the reference's atom order.  Units are GEMS internal units
(Angstrom, amu, ps, ui = amu A^2/ps^2).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_GEMS = os.path.join(HERE, "_ref", "gems")
MAXT = 8

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _setup(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Rng(C.Structure):
    _fields_ = [("idum", C.c_int32), ("ix", C.c_int32), ("iy", C.c_int32),
                ("stored", C.c_int), ("g", C.c_double)]


class Box(C.Structure):
    _fields_ = [("boxed", C.c_int), ("mic", C.c_int), ("pbc", C.c_int * 3),
                ("L", C.c_double * 3), ("one_box", C.c_double * 3)]


_MM = MAXT * MAXT


class Smatb(C.Structure):
    _fields_ = [("nt", C.c_int), ("prm", C.c_int * _MM)] + [
        (nm, C.c_double * _MM) for nm in
        ("ca", "eps", "p", "q", "r0", "rci", "rce", "b1", "b2", "b3", "r1", "r2", "r3")]


class Ermak(C.Structure):
    _fields_ = [(nm, C.c_double) for nm in
                ("c0", "c1", "c2", "sdr", "sdv", "crv1", "crv2", "skt", "dt", "gamma")]


class Plj(C.Structure):
    _fields_ = [("nt", C.c_int), ("kind", C.c_int), ("rc", C.c_double)] + [
        (nm, C.c_double * _MM) for nm in ("e", "s", "prc", "frc", "rs", "sa4", "sa8", "rcp")]


class MD(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("x", _dp), ("v", _dp), ("a", _dp), ("f", _dp), ("e", _dp),
        ("type", _ip), ("mass", _dp), ("box", Box),
        ("pot", C.c_int), ("tb", C.POINTER(Smatb)),
        ("lj_eps", C.c_double), ("lj_sigma", C.c_double), ("lj_rc", C.c_double),
        ("plj", C.c_void_p),
        ("skin", C.c_double),
        ("integ", C.c_int), ("dt", C.c_double), ("T", C.c_double), ("gamma", C.c_double),
        ("rng", C.POINTER(Rng)),
        ("nsteps", C.c_int), ("each", C.c_int), ("energy", _dp), ("nthreads", C.c_int),
        ("first_call", C.c_int), ("pos_old", _dp), ("maxnb", C.c_int),
        ("nn", _ip), ("list", _ip), ("boxcr", _ip), ("pos_v", _dp), ("vel_v", _dp),
        ("nupd", C.c_int), ("epot", C.c_double)]


def _setup(L):
    L.or_constants.argtypes = [_dp]
    L.or_element_mass.restype = C.c_double
    L.or_element_mass.argtypes = [C.c_int]
    L.or_rng_init.argtypes = [C.POINTER(Rng), C.c_int]
    L.or_ranu.restype = C.c_double
    L.or_ranu.argtypes = [C.POINTER(Rng)]
    L.or_gasdev.restype = C.c_double
    L.or_gasdev.argtypes = [C.POINTER(Rng)]
    L.or_box_set.argtypes = [C.POINTER(Box), C.c_int, _dp, _ip]
    L.or_nb_build.restype = C.c_int
    L.or_nb_build.argtypes = [_dp, C.c_int, C.POINTER(Box), C.c_double, C.c_int,
                              _ip, _ip, _ip, _ip, C.c_int]
    L.or_nb_brute.restype = C.c_int
    L.or_nb_brute.argtypes = [_dp, C.c_int, C.POINTER(Box), C.c_double, C.c_int,
                              C.c_int, _ip, _ip]
    L.or_smatb_init.argtypes = [C.POINTER(Smatb), C.c_int]
    L.or_smatb_set.argtypes = [C.POINTER(Smatb), C.c_int, C.c_int] + [C.c_double] * 7
    L.or_smatb_rcut.restype = C.c_double
    L.or_smatb_rcut.argtypes = [C.POINTER(Smatb)]
    L.or_smatb_force.argtypes = [_dp, _ip, C.c_int, C.POINTER(Box), C.POINTER(Smatb),
                                 _ip, _ip, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int]
    L.or_lj_force.argtypes = [_dp, C.c_int, C.POINTER(Box), C.c_double, C.c_double,
                              C.c_double, _ip, _ip, C.c_int, _dp, _dp, _dp, C.c_int]
    L.or_plj_init.argtypes = [C.POINTER(Plj), C.c_int, C.c_double]
    L.or_plj_set.argtypes = [C.POINTER(Plj), C.c_int, C.c_int, C.c_double, C.c_double]
    L.or_slj_set.argtypes = [C.POINTER(Plj), C.c_int, C.c_int, C.c_double, C.c_double,
                             C.c_double]
    L.or_wca_set.argtypes = [C.POINTER(Plj), C.c_int, C.c_int, C.c_double, C.c_double]
    L.or_sm1_set.argtypes = [C.POINTER(Plj), C.c_double, C.c_double, C.c_double]
    L.or_plj_force.argtypes = [_dp, _ip, C.c_int, C.POINTER(Box), C.POINTER(Plj), _ip, _ip,
                               C.c_int, _dp, _dp, _dp, C.c_int]
    L.or_ermak_set.argtypes = [C.POINTER(Ermak), C.c_double, C.c_double, C.c_double]
    L.or_ekin.restype = C.c_double
    L.or_ekin.argtypes = [C.c_int, _dp, _dp]
    L.or_set_gdist.argtypes = [C.POINTER(Rng), C.c_int, C.c_double, _dp, _dp, _dp]
    L.or_md_run.restype = C.c_int
    L.or_md_run.argtypes = [C.POINTER(MD)]
    L.or_md_interact.restype = C.c_int
    L.or_md_interact.argtypes = [C.POINTER(MD)]
    L.or_fire.restype = C.c_int
    L.or_fire.argtypes = [_dp, _ip, C.c_int, C.POINTER(Smatb), _dp, C.c_double, C.c_int,
                          C.c_double, C.c_double, C.c_double, _dp, _dp, _ip]


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def constants():
    out = np.zeros(5)
    lib().or_constants(_p(out))
    return dict(ev_ui=out[0], ui_ev=out[1], kB_ui=out[2], ui_kcm=out[3], kB_ev=out[4])


def element_mass(z):
    return lib().or_element_mass(z)


class LCG:
    """The reference's ``prng lcg`` stream (Random.F90:145-250)."""

    def __init__(self, seed):
        self.s = Rng()
        lib().or_rng_init(C.byref(self.s), int(seed))

    def ranu(self):
        return lib().or_ranu(C.byref(self.s))

    def gasdev(self):
        return lib().or_gasdev(C.byref(self.s))


def make_box(L=None, pbc=(0, 0, 0)):
    b = Box()
    if L is None:
        lib().or_box_set(C.byref(b), 0, None, None)
    else:
        Ls = np.ascontiguousarray(L, dtype=np.float64)
        pb = np.ascontiguousarray(pbc, dtype=np.int32)
        lib().or_box_set(C.byref(b), 1, _p(Ls), _p(pb, _ip))
    return b


def smatb_params(rows, nt):
    """rows: list of (i, j, A, xi, p, q, r0, rci, rce) with 0-based types."""
    s = Smatb()
    lib().or_smatb_init(C.byref(s), nt)
    for (i, j, A, xi, p, q, r0, rci, rce) in rows:
        lib().or_smatb_set(C.byref(s), i, j, A, xi, p, q, r0, rci, rce)
    return s


def smatb_rcut(s):
    return lib().or_smatb_rcut(C.byref(s))


def neighbours(x, box, rlist, maxnb=256, nthreads=1):
    """Reference list (verlet_self or linkedcell_self): returns (nn, list, mode)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    while True:
        nn = np.zeros(n, dtype=np.int32)
        lst = np.zeros((n, maxnb), dtype=np.int32)
        mode = C.c_int(0)
        nc = (C.c_int * 3)()
        rc = lib().or_nb_build(_p(x), n, C.byref(box), rlist, maxnb, _p(nn, _ip),
                               _p(lst, _ip), C.byref(mode), nc, nthreads)
        if rc > 0:
            maxnb = rc + 8
            continue
        return nn, lst, mode.value


def neighbours_brute(x, box, rlist, strict=True, maxnb=256):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    while True:
        nn = np.zeros(n, dtype=np.int32)
        lst = np.zeros((n, maxnb), dtype=np.int32)
        rc = lib().or_nb_brute(_p(x), n, C.byref(box), rlist, int(strict), maxnb,
                               _p(nn, _ip), _p(lst, _ip))
        if rc > 0:
            maxnb = rc + 8
            continue
        return nn, lst


def sorted_sets(nn, lst):
    return [np.sort(lst[i, :nn[i]]) for i in range(len(nn))]


def smatb_force(x, types, box, s, nn, lst, nthreads=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    t = np.ascontiguousarray(types, dtype=np.int32)
    f = np.zeros((n, 3))
    e = np.zeros(n)
    rho = np.zeros(n)
    eb = np.zeros(n)
    ep = C.c_double(0)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    lib().or_smatb_force(_p(x), _p(t, _ip), n, C.byref(box), C.byref(s), _p(nn, _ip),
                         _p(lst, _ip), lst.shape[1], _p(f), _p(e), _p(rho), _p(eb),
                         C.byref(ep), nthreads)
    return f, e, ep.value, rho


def lj_force(x, box, eps, sigma, rc, nn, lst, nthreads=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    f = np.zeros((n, 3))
    e = np.zeros(n)
    ep = C.c_double(0)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    lib().or_lj_force(_p(x), n, C.byref(box), eps, sigma, rc, _p(nn, _ip), _p(lst, _ip),
                      lst.shape[1], _p(f), _p(e), C.byref(ep), nthreads)
    return f, e, ep.value


def plj_params(eps, sigma, rc=10.0):
    """plj_set (Pairs.F90:282-291) over nt x nt symmetric eps (eV), sigma (A)."""
    eps = np.asarray(eps, dtype=np.float64)
    sigma = np.asarray(sigma, dtype=np.float64)
    nt = eps.shape[0]
    p = Plj()
    lib().or_plj_init(C.byref(p), nt, rc)
    for i in range(nt):
        for j in range(i, nt):
            lib().or_plj_set(C.byref(p), i, j, eps[i, j], sigma[i, j])
    return p


def slj_params(eps, sigma, rs):
    """slj_set (Pairs.F90:374-401) over nt x nt symmetric eps, sigma, rs."""
    eps, sigma, rs = (np.asarray(a, dtype=np.float64) for a in (eps, sigma, rs))
    nt = eps.shape[0]
    p = Plj()
    lib().or_plj_init(C.byref(p), nt, 0.0)
    p.kind = 1
    for i in range(nt):
        for j in range(i, nt):
            lib().or_slj_set(C.byref(p), i, j, eps[i, j], sigma[i, j], rs[i, j])
    return p


def wca_params(eps, sigma):
    """wca_set (Pairs.F90:495-506) over nt x nt symmetric eps, sigma."""
    eps, sigma = (np.asarray(a, dtype=np.float64) for a in (eps, sigma))
    nt = eps.shape[0]
    p = Plj()
    lib().or_plj_init(C.byref(p), nt, 0.0)
    p.kind = 2
    for i in range(nt):
        for j in range(i, nt):
            lib().or_wca_set(C.byref(p), i, j, eps[i, j], sigma[i, j])
    return p


def sm1_params(r1, r2, e, nt=1):
    """sm1_set (Pairs.F90:916-922): r1, cut r2 (A), e (eV) for every pair."""
    p = Plj()
    lib().or_plj_init(C.byref(p), nt, 0.0)
    lib().or_sm1_set(C.byref(p), r1, r2, e)
    return p


def plj_force(x, types, box, p, nn, lst, nthreads=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    t = np.ascontiguousarray(types, dtype=np.int32)
    f = np.zeros((n, 3))
    e = np.zeros(n)
    ep = C.c_double(0)
    lst = np.ascontiguousarray(lst, dtype=np.int32)
    lib().or_plj_force(_p(x), _p(t, _ip), n, C.byref(box), C.byref(p), _p(nn, _ip),
                       _p(lst, _ip), lst.shape[1], _p(f), _p(e), C.byref(ep), nthreads)
    return f, e, ep.value


def ekin(v, mass):
    v = np.ascontiguousarray(v, dtype=np.float64)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    return lib().or_ekin(len(m), _p(v), _p(m))


def set_gdist(rng, T, mass, v=None):
    m = np.ascontiguousarray(mass, dtype=np.float64)
    n = len(m)
    v = np.zeros((n, 3)) if v is None else np.ascontiguousarray(v, dtype=np.float64).copy()
    om = 1.0 / m
    lib().or_set_gdist(C.byref(rng.s), n, T, _p(v), _p(m), _p(om))
    return v


def ermak_coeffs(dt, gamma, T):
    k = Ermak()
    lib().or_ermak_set(C.byref(k), dt, gamma, T)
    return {nm: getattr(k, nm) for nm, _ in Ermak._fields_}


class MDRun:
    """The reference's ``dinamica`` loop on the CPU (Programs.f90:114-167)."""

    def __init__(self, x, v, mass, box, pot="smatb", tb=None, types=None, lj=None, plj=None,
                 skin=1.0, integ="vv", dt=0.002, T=0.0, gamma=1.0, rng=None,
                 nthreads=1, maxnb=128):
        n = x.shape[0]
        self.n = n
        self.x = np.ascontiguousarray(x, dtype=np.float64).copy()
        self.v = np.ascontiguousarray(v, dtype=np.float64).copy()
        self.a = np.zeros((n, 3))
        self.f = np.zeros((n, 3))
        self.e = np.zeros(n)
        self.mass = np.ascontiguousarray(mass, dtype=np.float64)
        self.types = np.ascontiguousarray(types if types is not None else np.zeros(n),
                                          dtype=np.int32)
        self.pos_old = np.zeros((n, 3))
        self.maxnb = maxnb
        self.nn = np.zeros(n, dtype=np.int32)
        self.list = np.zeros((n, maxnb), dtype=np.int32)
        self.boxcr = np.zeros((n, 3), dtype=np.int32)
        self.pos_v = np.zeros((n, 3))
        self.vel_v = np.zeros((n, 3))
        self.tb = tb
        self.rng = rng if rng is not None else LCG(123456)
        md = MD()
        md.n = n
        md.box = box
        md.pot = {"smatb": 0, "lj": 1, "plj": 2}[pot]
        if tb is not None:
            md.tb = C.pointer(tb)
        if lj is not None:
            md.lj_eps, md.lj_sigma, md.lj_rc = lj
        self.plj = plj
        if plj is not None:
            md.plj = C.cast(C.pointer(plj), C.c_void_p)
        md.skin = skin
        md.integ = 0 if integ == "vv" else 1
        md.dt, md.T, md.gamma = dt, T, gamma
        md.rng = C.pointer(self.rng.s)
        md.nthreads = nthreads
        md.first_call = 1
        self.md = md
        self._bind()

    def _bind(self):
        md = self.md
        md.x, md.v, md.a, md.f, md.e = (_p(self.x), _p(self.v), _p(self.a),
                                        _p(self.f), _p(self.e))
        md.type = _p(self.types, _ip)
        md.mass = _p(self.mass)
        md.pos_old = _p(self.pos_old)
        md.maxnb = self.maxnb
        md.nn = _p(self.nn, _ip)
        md.list = _p(self.list, _ip)
        md.boxcr = _p(self.boxcr, _ip)
        md.pos_v = _p(self.pos_v)
        md.vel_v = _p(self.vel_v)

    def _grow(self, need):
        self.maxnb = need + 16
        self.list = np.zeros((self.n, self.maxnb), dtype=np.int32)
        self.md.first_call = 1
        self.md.nupd -= 1
        self._bind()

    def interact(self):
        while True:
            rc = lib().or_md_interact(C.byref(self.md))
            if rc > 0:
                self._grow(rc)
                continue
            if rc < 0:
                raise RuntimeError("oracle interact failed")
            return self.md.epot

    def run(self, nsteps, each=1):
        nout = nsteps // each if each > 0 else 0
        en = np.zeros((max(nout, 1), 3))
        self.md.nsteps = nsteps
        self.md.each = each
        self.md.energy = _p(en)
        rc = lib().or_md_run(C.byref(self.md))
        if rc != 0:
            raise RuntimeError("oracle md_run failed (%d)" % rc)
        return en[:nout]

    @property
    def nupd(self):
        return self.md.nupd


def fire(x, types, s, masses, skin=1.0, max_iter=20000, dt_max=0.005, alpha0=0.1, ftol=1e-4):
    """FP64 FIRE relaxation of one free system (see gems_oracle.h)."""
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    t = np.ascontiguousarray(types, dtype=np.int32)
    m = np.ascontiguousarray(masses, dtype=np.float64)
    ep, fm, cv = C.c_double(), C.c_double(), C.c_int()
    it = lib().or_fire(_p(x), _p(t, _ip), len(t), C.byref(s), _p(m), skin, max_iter, dt_max,
                       alpha0, ftol, C.byref(ep), C.byref(fm), C.byref(cv))
    return dict(x=x, iters=it, epot=ep.value, fmax=fm.value, converged=bool(cv.value))
