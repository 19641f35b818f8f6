"""ctypes binding of the gfx950 library (include/gems_hip.h).

The library is loaded from gems_amd/lib/libgems_hip.so, built in-tree by
``python gems_amd/build.py`` (or ``__graft_entry__.build()``).  There is no
fallback: if the library is missing, importing this module raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgems_hip.so")

GH_FP64, GH_MIXED = 0, 1
GH_RNG_LCG, GH_RNG_PHILOX = 0, 1
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ESTATE", -4: "EOVERFLOW", -5: "ERANGE"}

# every symbol declared in include/gems_hip.h
SYMBOLS = [
    "gh_create", "gh_destroy", "gh_last_error", "gh_set_precision", "gh_set_box", "gh_set_images", "gh_set_pair_masks", "gh_set_sm1",
    "gh_set_atoms", "gh_set_positions", "gh_set_masses", "gh_set_velocities",
    "gh_set_accelerations", "gh_set_smatb",
    "gh_set_lj", "gh_set_plj", "gh_set_slj", "gh_set_wca", "gh_get_lcg_state",
    "gh_set_lcg_state", "gh_lbfgs",
    "gh_set_skin", "gh_compute", "gh_neighbors", "gh_prng_lcg", "gh_set_gdist",
    "gh_integrator_vv", "gh_integrator_ermak", "gh_run", "gh_energies", "gh_get_state",
    "gh_get_stats", "gh_set_timing", "gh_get_timing", "gh_set_tile", "gh_autotune",
    "gh_batch_create", "gh_batch_destroy", "gh_batch_last_error", "gh_batch_set_precision",
    "gh_batch_set_systems", "gh_batch_set_smatb", "gh_batch_set_masses", "gh_batch_set_skin",
    "gh_batch_set_velocities", "gh_batch_set_temperatures", "gh_batch_compute",
    "gh_batch_langevin", "gh_batch_exchange", "gh_batch_fire", "gh_batch_get_state",
    "gh_cluster_forces_static",
]


class GhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Stats(C.Structure):
    _fields_ = [("nupd", C.c_int64), ("maxnn", C.c_int32), ("cap", C.c_int32),
                ("ncells", C.c_int32 * 3), ("list_mode", C.c_int32), ("steps", C.c_int64),
                ("launches", C.c_int64), ("list_build", C.c_int32),
                ("chunks", C.c_int32)]


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def load():
    """Loads the in-tree library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("gems_amd: %s is missing; build it with "
                          "`python gems_amd/build.py`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.gh_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int]
    L.gh_destroy.argtypes = [vp]
    L.gh_destroy.restype = None
    L.gh_last_error.argtypes = [vp]
    L.gh_last_error.restype = C.c_char_p
    L.gh_set_precision.argtypes = [vp, C.c_int]
    L.gh_set_box.argtypes = [vp, _dp, _ip]
    L.gh_set_images.argtypes = [vp, C.c_int]
    L.gh_set_pair_masks.argtypes = [vp, _ip, _ip]
    L.gh_set_sm1.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    L.gh_set_atoms.argtypes = [vp, C.c_int, _dp, _ip]
    L.gh_set_masses.argtypes = [vp, C.c_int, _dp]
    L.gh_set_velocities.argtypes = [vp, _dp]
    L.gh_set_accelerations.argtypes = [vp, _dp]
    L.gh_set_positions.argtypes = [vp, _dp]
    L.gh_set_smatb.argtypes = [vp, C.c_int] + [_dp] * 7
    L.gh_set_lj.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    L.gh_set_plj.argtypes = [vp, C.c_int, _dp, _dp, C.c_double]
    L.gh_set_slj.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    L.gh_set_wca.argtypes = [vp, C.c_int, _dp, _dp]
    L.gh_set_skin.argtypes = [vp, C.c_double]
    L.gh_set_tile.argtypes = [vp, _ip]
    L.gh_autotune.argtypes = [vp, C.c_int, _ip, _dp]
    L.gh_compute.argtypes = [vp, _dp, _dp, _dp]
    L.gh_neighbors.argtypes = [vp, _ip, _ip, C.c_int]
    L.gh_prng_lcg.argtypes = [vp, C.c_int]
    L.gh_get_lcg_state.argtypes = [vp, _ip]
    L.gh_lbfgs.argtypes = [vp, C.c_int, C.c_double, C.c_int, _ip, _dp]
    L.gh_set_lcg_state.argtypes = [vp, _ip]
    L.gh_set_gdist.argtypes = [vp, C.c_double]
    L.gh_integrator_vv.argtypes = [vp, C.c_double]
    L.gh_integrator_ermak.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_uint64]
    L.gh_run.argtypes = [vp, C.c_int, C.c_int, _dp]
    L.gh_energies.argtypes = [vp, _dp, _dp]
    L.gh_get_state.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.gh_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.gh_set_timing.argtypes = [vp, C.c_int]
    L.gh_get_timing.argtypes = [vp, _dp, C.POINTER(C.c_int64)]
    L.gh_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int]
    L.gh_batch_destroy.argtypes = [vp]
    L.gh_batch_destroy.restype = None
    L.gh_batch_last_error.argtypes = [vp]
    L.gh_batch_last_error.restype = C.c_char_p
    L.gh_batch_set_precision.argtypes = [vp, C.c_int]
    L.gh_batch_set_systems.argtypes = [vp, C.c_int, _ip, _dp, _ip]
    L.gh_batch_set_smatb.argtypes = [vp, C.c_int] + [_dp] * 7
    L.gh_batch_set_masses.argtypes = [vp, C.c_int, _dp]
    L.gh_batch_set_skin.argtypes = [vp, C.c_double]
    L.gh_batch_set_velocities.argtypes = [vp, _dp]
    L.gh_batch_set_temperatures.argtypes = [vp, _dp]
    L.gh_batch_compute.argtypes = [vp, _dp, _dp]
    L.gh_batch_langevin.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_uint64, _dp, _dp]
    L.gh_batch_exchange.argtypes = [vp, C.c_int, C.c_int, _ip, C.c_int, C.c_uint64, _ip]
    L.gh_batch_fire.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                _ip, _ip, _dp, _dp]
    L.gh_batch_get_state.argtypes = [vp, _dp, _dp]
    L.gh_cluster_forces_static.argtypes = ([C.c_int] * 4 + [vp, vp] + [_dp] * 7 +
                                           [C.c_double] + [vp] * 5)
    _lib = L
    return L


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


class Context:
    """One system on one device (a gh_ctx)."""

    def __init__(self, device=0, precision=GH_FP64):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.gh_create(C.byref(h), device, precision)
        if rc != 0:
            raise GhError(rc, "gh_create failed (no HIP device %d?)" % device)
        self.h = h
        self.n = 0

    def _ck(self, rc):
        if rc != 0:
            raise GhError(rc, self.L.gh_last_error(self.h).decode())

    def close(self):
        if self.h is not None:
            self.L.gh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- setup ---
    def set_precision(self, p):
        self._ck(self.L.gh_set_precision(self.h, p))

    def set_box(self, L=None, pbc=(1, 1, 1)):
        if L is None:
            self._ck(self.L.gh_set_box(self.h, None, None))
        else:
            Ls = np.ascontiguousarray(L, np.float64)
            pb = np.ascontiguousarray(pbc, np.int32)
            self._ck(self.L.gh_set_box(self.h, _p(Ls), _p(pb, _ip)))

    def set_images(self, ghost):
        """`box mic F` (ghost=True): every first-shell periodic image of every
        atom, as the reference's ghost atoms; `box mic T` (False, the
        default): the nearest image only."""
        self._ck(self.L.gh_set_images(self.h, 1 if ghost else 0))

    def set_pair_masks(self, feel=None, felt=None):
        """One-sided interactions (`interact A feels B`): atom i pairs with j
        only when feel[i] & felt[j] != 0; None, None: every pair."""
        if feel is None:
            self._ck(self.L.gh_set_pair_masks(self.h, None, None))
            return
        fe = np.ascontiguousarray(feel, np.int32)
        ft = np.ascontiguousarray(felt, np.int32)
        self._ck(self.L.gh_set_pair_masks(self.h, _p(fe, _ip), _p(ft, _ip)))

    def set_atoms(self, x, types=None):
        x = np.ascontiguousarray(x, np.float64)
        self.n = x.shape[0]
        t = None if types is None else np.ascontiguousarray(types, np.int32)
        self._ck(self.L.gh_set_atoms(self.h, self.n, _p(x), _p(t, _ip)))

    def set_masses(self, masses):
        m = np.ascontiguousarray(np.atleast_1d(masses), np.float64)
        self._ck(self.L.gh_set_masses(self.h, len(m), _p(m)))

    def set_velocities(self, v=None):
        vv = None if v is None else np.ascontiguousarray(v, np.float64)
        self._ck(self.L.gh_set_velocities(self.h, _p(vv)))

    def set_accelerations(self, a=None):
        aa = None if a is None else np.ascontiguousarray(a, np.float64)
        self._ck(self.L.gh_set_accelerations(self.h, _p(aa)))

    def set_positions(self, x):
        """New positions (n x 3, caller order) for the next compute; lists kept."""
        xx = np.ascontiguousarray(x, np.float64)
        assert xx.shape == (self.n, 3)
        self._ck(self.L.gh_set_positions(self.h, _p(xx)))

    def set_smatb(self, nt, A, xi, p, q, r0, rci, rce):
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (nt, nt)))
                for a in (A, xi, p, q, r0, rci, rce)]
        self._ck(self.L.gh_set_smatb(self.h, nt, *[_p(a) for a in arrs]))

    def set_lj(self, eps, sigma, rc):
        self._ck(self.L.gh_set_lj(self.h, eps, sigma, rc))

    def set_plj(self, eps, sigma, rc=10.0):
        """Shifted Lennard-Jones (plj, Pairs.F90:282-370): eps (eV) and sigma
        (A) are nt x nt symmetric matrices over the atom types; rc in A."""
        e = np.ascontiguousarray(eps, dtype=np.float64)
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        nt = e.shape[0]
        self._ck(self.L.gh_set_plj(self.h, nt, e.ctypes.data_as(_dp),
                                   sg.ctypes.data_as(_dp), rc))

    def set_slj(self, eps, sigma, rs):
        """Splined Lennard-Jones (slj, Pairs.F90:374-493): nt x nt eps (eV),
        sigma and rs (A); the cut follows from them."""
        e, sg, r = (np.ascontiguousarray(a, dtype=np.float64) for a in (eps, sigma, rs))
        self._ck(self.L.gh_set_slj(self.h, e.shape[0], e.ctypes.data_as(_dp),
                                   sg.ctypes.data_as(_dp), r.ctypes.data_as(_dp)))

    def set_sm1(self, r1, r2, e):
        """sm1 smooth step (Pairs.F90:916-990): -e inside r1, a sine ramp to
        zero at the cut r2 (A, eV)."""
        self._ck(self.L.gh_set_sm1(self.h, r1, r2, e))

    def set_wca(self, eps, sigma):
        """Weeks-Chandler-Andersen (wca, Pairs.F90:495-560): nt x nt eps, sigma."""
        e, sg = (np.ascontiguousarray(a, dtype=np.float64) for a in (eps, sigma))
        self._ck(self.L.gh_set_wca(self.h, e.shape[0], e.ctypes.data_as(_dp),
                                   sg.ctypes.data_as(_dp)))

    def set_skin(self, skin):
        self._ck(self.L.gh_set_skin(self.h, skin))

    def set_tile(self, bd=None):
        """Brick edges (cells) of the tiled passes; None = automatic."""
        b = None if bd is None else np.ascontiguousarray(bd, np.int32)
        self._ck(self.L.gh_set_tile(self.h, _p(b, _ip)))

    def autotune(self, reps=5):
        """Times the force passes per brick shape and keeps the fastest;
        returns (shape, microseconds per force evaluation)."""
        bd = np.zeros(3, np.int32)
        us = C.c_double(0.0)
        self._ck(self.L.gh_autotune(self.h, reps, _p(bd, _ip), C.byref(us)))
        return tuple(int(v) for v in bd), us.value

    # --- evaluation ---
    def compute(self):
        F = np.zeros((self.n, 3))
        e = np.zeros(self.n)
        et = C.c_double(0.0)
        self._ck(self.L.gh_compute(self.h, _p(F), _p(e), C.byref(et)))
        return F, e, et.value

    def neighbors(self, cap=256):
        nn = np.zeros(self.n, np.int32)
        idx = np.zeros((self.n, cap), np.int32)
        self._ck(self.L.gh_neighbors(self.h, _p(nn, _ip), _p(idx, _ip), cap))
        return nn, idx

    def prng_lcg(self, seed):
        self._ck(self.L.gh_prng_lcg(self.h, seed))

    def lbfgs(self, m=5, eps=1e-8, max_iter=10000):
        """`lbfgs` (Quasi_Newton.f90:160-232) on the device force path; returns
        (iterations, energy in eV); the minimum stays in the context."""
        it = C.c_int(0)
        ep = C.c_double(0.0)
        self._ck(self.L.gh_lbfgs(self.h, m, eps, max_iter, C.byref(it), C.byref(ep)))
        return it.value, ep.value

    def lcg_state(self, state=None):
        """The lcg stream's (idum, iy, ix) as a checkpoint stores it; with
        `state`, sets it first."""
        st = np.zeros(3, np.int32)
        if state is not None:
            st[:] = state
            self._ck(self.L.gh_set_lcg_state(self.h, _p(st, _ip)))
        self._ck(self.L.gh_get_lcg_state(self.h, _p(st, _ip)))
        return tuple(int(v) for v in st)

    def set_gdist(self, T):
        self._ck(self.L.gh_set_gdist(self.h, T))

    def integrator_vv(self, dt):
        self._ck(self.L.gh_integrator_vv(self.h, dt))

    def integrator_ermak(self, dt, T, gamma, rng=GH_RNG_LCG, seed=0):
        self._ck(self.L.gh_integrator_ermak(self.h, dt, T, gamma, rng, seed))

    def run(self, nsteps, each=0):
        nout = nsteps // each if each > 0 else 0
        en = np.zeros((max(nout, 1), 3))
        self._ck(self.L.gh_run(self.h, nsteps, each, _p(en) if each > 0 else None))
        return en[:nout]

    def energies(self):
        ep, ek = C.c_double(), C.c_double()
        self._ck(self.L.gh_energies(self.h, C.byref(ep), C.byref(ek)))
        return ep.value, ek.value

    def state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        F = np.zeros((self.n, 3))
        e = np.zeros(self.n)
        self._ck(self.L.gh_get_state(self.h, _p(x), _p(v), _p(F), _p(e)))
        return x, v, F, e

    def stats(self):
        s = Stats()
        self._ck(self.L.gh_get_stats(self.h, C.byref(s)))
        return dict(nupd=s.nupd, maxnn=s.maxnn, cap=s.cap, ncells=tuple(s.ncells),
                    list_mode=s.list_mode, steps=s.steps, launches=s.launches,
                    list_build=s.list_build, chunks=s.chunks)

    def set_timing(self, on=True):
        self._ck(self.L.gh_set_timing(self.h, int(on)))

    def timing(self):
        ms = np.zeros(4)
        cnt = (C.c_int64 * 4)()
        self._ck(self.L.gh_get_timing(self.h, _p(ms), cnt))
        names = ["rho", "force", "integrate", "rebuild"]
        return {k: (ms[i], cnt[i]) for i, k in enumerate(names)}


class Batch:
    """Many independent free clusters on one device (a gh_batch)."""

    def __init__(self, device=0, precision=GH_FP64):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.gh_batch_create(C.byref(h), device, precision)
        if rc != 0:
            raise GhError(rc, "gh_batch_create failed (no HIP device %d?)" % device)
        self.h = h
        self.nsys = 0
        self.ntot = 0

    def _ck(self, rc):
        if rc != 0:
            raise GhError(rc, self.L.gh_batch_last_error(self.h).decode())

    def close(self):
        if self.h is not None:
            self.L.gh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_precision(self, p):
        self._ck(self.L.gh_batch_set_precision(self.h, p))

    def set_systems(self, off, x, types=None):
        off = np.ascontiguousarray(off, np.int32)
        x = np.ascontiguousarray(x, np.float64)
        self.nsys = len(off) - 1
        self.ntot = int(off[-1])
        self.off = off
        t = None if types is None else np.ascontiguousarray(types, np.int32)
        self._ck(self.L.gh_batch_set_systems(self.h, self.nsys, _p(off, _ip), _p(x),
                                             _p(t, _ip)))

    def set_smatb(self, nt, A, xi, p, q, r0, rci, rce):
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (nt, nt)))
                for a in (A, xi, p, q, r0, rci, rce)]
        self._ck(self.L.gh_batch_set_smatb(self.h, nt, *[_p(a) for a in arrs]))

    def set_masses(self, masses):
        m = np.ascontiguousarray(np.atleast_1d(masses), np.float64)
        self._ck(self.L.gh_batch_set_masses(self.h, len(m), _p(m)))

    def set_skin(self, skin):
        self._ck(self.L.gh_batch_set_skin(self.h, skin))

    def set_velocities(self, v=None):
        vv = None if v is None else np.ascontiguousarray(v, np.float64)
        self._ck(self.L.gh_batch_set_velocities(self.h, _p(vv)))

    def set_temperatures(self, T):
        T = np.ascontiguousarray(T, np.float64)
        self._ck(self.L.gh_batch_set_temperatures(self.h, _p(T)))

    def compute(self):
        F = np.zeros((self.ntot, 3))
        ep = np.zeros(self.nsys)
        self._ck(self.L.gh_batch_compute(self.h, _p(F), _p(ep)))
        return F, ep

    def langevin(self, nsteps, dt, gamma, seed):
        ep = np.zeros(self.nsys)
        ek = np.zeros(self.nsys)
        self._ck(self.L.gh_batch_langevin(self.h, nsteps, dt, gamma, seed, _p(ep), _p(ek)))
        return ep, ek

    def exchange(self, ladders, parity, seed):
        """ladders: (nladder, R) int array of system ids in rung order (updated)."""
        lad = np.ascontiguousarray(ladders, np.int32)
        nacc = C.c_int(0)
        self._ck(self.L.gh_batch_exchange(self.h, lad.shape[0], lad.shape[1], _p(lad, _ip),
                                          parity, seed, C.byref(nacc)))
        ladders[...] = lad
        return nacc.value

    def fire(self, max_iter=20000, dt_max=0.005, alpha0=0.1, fswitch=1e-2, ftol=1e-4):
        it = np.zeros(self.nsys, np.int32)
        cv = np.zeros(self.nsys, np.int32)
        ep = np.zeros(self.nsys)
        fm = np.zeros(self.nsys)
        self._ck(self.L.gh_batch_fire(self.h, max_iter, dt_max, alpha0, fswitch, ftol,
                                      _p(it, _ip), _p(cv, _ip), _p(ep), _p(fm)))
        return dict(iters=it, converged=cv, epot=ep, fmax=fm)

    def state(self):
        x = np.zeros((self.ntot, 3))
        v = np.zeros((self.ntot, 3))
        self._ck(self.L.gh_batch_get_state(self.h, _p(x), _p(v)))
        return x, v
