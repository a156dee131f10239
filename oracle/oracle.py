"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py`` as the checker.  The product
(``parm_amd`` and ``libparm_hip.so``) never imports or calls it.

The oracle restates the reference's double-precision CPU path; see
``parm_oracle.h`` for the citations (src/box.hpp, src/trackers.cpp,
src/interaction.hpp, src/collection.cpp of wackywendell/parm).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

LJ_ATTRACT_REPULSE = 1
REPULSION = 2
LJ_CUT = 3           # NListed<EpsSigCutAtom, LennardJonesCutPair>
LJ_CUT_II = 4        # NListed<IEpsISigCutAtom, LennardJonesCutPair>
LJ_REPULSE = 5       # NListed<EpsSigAtom, LJRepulsePair>
REPULSION_II = 6     # NListed<IEpsISigExpAtom, RepulsionPair>
_EPS_ROWS = (LJ_ATTRACT_REPULSE, LJ_CUT_II, REPULSION_II)
_SIG_ROWS = (LJ_CUT_II, REPULSION_II)

_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint64)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_vec_mod1.restype = C.c_double
        L.or_vec_mod1.argtypes = [C.c_double, C.c_double]
        L.or_pair.restype = C.c_double
        L.or_pair.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int,
                              C.c_int, _dp, _dp, _dp, _dp]
        for nm in ("or_nlist_brute", "or_nlist_cells"):
            fn = getattr(L, nm)
            fn.restype = C.c_int64
            fn.argtypes = [C.c_int, C.c_int64, _ip, _dp, _dp, _dp, C.c_double,
                           C.c_int64, _ip, _ip, C.c_int64, _ip, _ip]
        L.or_skin_check.restype = C.c_int
        L.or_skin_check.argtypes = [C.c_int, C.c_int64, _ip, _dp, _dp, C.c_double]
        L.or_forces.restype = None
        L.or_forces.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int, C.c_int,
                                _dp, _dp, _dp, _ip, C.c_int64, _ip, _ip, _dp,
                                _dp, _dp, _dp, _dp, _up, _up]
        L.or_sys_create.restype = C.c_void_p
        L.or_sys_create.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int, C.c_int,
                                    _dp, _dp, _dp, _ip, C.c_int64, _ip, _dp,
                                    C.c_double, C.c_int]
        L.or_sys_destroy.argtypes = [C.c_void_p]
        L.or_sys_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.or_sys_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.or_sys_ignore.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.or_sys_ignore_size.restype = C.c_int64
        L.or_sys_ignore_size.argtypes = [C.c_void_p]
        L.or_sys_update_list.restype = C.c_int
        L.or_sys_update_list.argtypes = [C.c_void_p, C.c_int]
        L.or_sys_which.restype = C.c_uint32
        L.or_sys_which.argtypes = [C.c_void_p]
        L.or_sys_numpairs.restype = C.c_int64
        L.or_sys_numpairs.argtypes = [C.c_void_p]
        L.or_sys_get_pairs.restype = C.c_int64
        L.or_sys_get_pairs.argtypes = [C.c_void_p, C.c_int64, _ip, _ip]
        L.or_sys_set_forces.argtypes = [C.c_void_p, C.c_int]
        for nm in ("or_sys_potential_energy", "or_sys_virial", "or_sys_mass"):
            getattr(L, nm).restype = C.c_double
            getattr(L, nm).argtypes = [C.c_void_p]
        L.or_sys_stress.argtypes = [C.c_void_p, _dp]
        L.or_sys_kinetic_energy.restype = C.c_double
        L.or_sys_kinetic_energy.argtypes = [C.c_void_p, _dp]
        L.or_sys_momentum.argtypes = [C.c_void_p, _dp]
        L.or_sys_temp.restype = C.c_double
        L.or_sys_temp.argtypes = [C.c_void_p, C.c_int]
        L.or_sys_initialize.argtypes = [C.c_void_p]
        L.or_sys_verlet.argtypes = [C.c_void_p, C.c_double, C.c_int64]
        L.or_damped_constants.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp]
        L.or_sys_damped.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int64]
        L.or_nlcg_create.restype = C.c_void_p
        L.or_nlcg_create.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int, C.c_double, C.c_int]
        L.or_nlcg_destroy.argtypes = [C.c_void_p]
        L.or_nlcg_set_limits.argtypes = [C.c_void_p] + [C.c_double] * 5
        L.or_nlcg_set_forces.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.or_nlcg_reset.argtypes = [C.c_void_p, C.c_void_p]
        L.or_nlcg_timestep.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.or_nlcg_descend.argtypes = [C.c_void_p, C.c_void_p]
        L.or_nlcg_observables.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, _dp]
        L.or_nlcg_info.argtypes = [C.c_void_p, _dp]
        L.or_sys_box.argtypes = [C.c_void_p, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def vec_mod(dx: float, L: float) -> float:
    return lib().or_vec_mod1(float(dx), float(L))


class PairTable:
    """Per-type parameters of one pair interaction, as the reference's per-atom
    structs hold them.  eps (and sigma) are per type, or ntypes x ntypes rows
    for the indexed atom types (IEpsSigCutAtom's epsilons; IEpsISigCutAtom's
    and IEpsISigExpAtom's epsilons and sigmas).  third is sigcut or the
    exponent per type (ignored by LJ_REPULSE)."""

    def __init__(self, kind, eps, sigma, third, types):
        self.kind = int(kind)
        sig = _f64(sigma)
        nt = int(round(sig.size ** 0.5)) if self.kind in _SIG_ROWS else sig.size
        self.sigma = _f64(sig, (nt, nt)) if self.kind in _SIG_ROWS else sig.ravel()
        if self.kind in _EPS_ROWS:
            self.eps = _f64(eps, (nt, nt))
        else:
            self.eps = _f64(eps).ravel()
        self.third = (np.zeros(nt) if third is None else _f64(third).ravel())
        self.types = _i32(types)
        self.ntypes = nt


def pair(ndim, table: PairTable, t1, t2, r1, r2, L):
    f = np.zeros(ndim)
    E = lib().or_pair(ndim, table.kind, table.ntypes, _d(table.eps), _d(table.sigma),
                      _d(table.third), int(t1), int(t2), _d(_f64(r1)), _d(_f64(r2)),
                      _d(_f64(L)), _d(f))
    return E, f


def _members(natoms, member):
    return _i32(np.arange(natoms) if member is None else member)


def nlist(x, L, diam, skin, member=None, brute=False, ignore=None):
    """Reference-order pair list (first, last) as int32 arrays.  ignore: an
    array of excluded (a, b) atom-index pairs (NeighborList::ignore)."""
    x = _f64(x)
    n, ndim = x.shape
    mem = _members(n, member)
    diam = _f64(diam).ravel()
    L = _f64(L).ravel()
    ex = np.zeros((0, 2), np.int32) if ignore is None else _i32(ignore).reshape(-1, 2)
    ea, eb = _i32(ex[:, 0]), _i32(ex[:, 1])
    fn = lib().or_nlist_brute if brute else lib().or_nlist_cells
    cap = max(16, 64 * len(mem))
    while True:
        first = np.empty(cap, np.int32)
        last = np.empty(cap, np.int32)
        npairs = fn(ndim, len(mem), _i(mem), _d(diam), _d(x), _d(L), float(skin), cap,
                    _i(first), _i(last), len(ex), _i(ea), _i(eb))
        if npairs <= cap:
            return first[:npairs].copy(), last[:npairs].copy()
        cap = int(npairs)


def skin_check(x, lastlocs, skin, member=None):
    x = _f64(x)
    n, ndim = x.shape
    mem = _members(n, member)
    return bool(lib().or_skin_check(ndim, len(mem), _i(mem), _d(x),
                                    _d(_f64(lastlocs)), float(skin)))


def forces(x, L, table: PairTable, first, last):
    """NListed::set_forces + energy + pressure + stress over a pair list.

    Returns dict(f, fabs, U, W, S, contacts, overlaps)."""
    x = _f64(x)
    n, ndim = x.shape
    first = _i32(first)
    last = _i32(last)
    f = np.zeros((n, ndim))
    fabs = np.zeros(n)
    U = np.zeros(1)
    W = np.zeros(1)
    S = np.zeros((ndim, ndim))
    nc = np.zeros(1, np.uint64)
    no = np.zeros(1, np.uint64)
    lib().or_forces(ndim, n, _d(x), _d(_f64(L)), table.kind, table.ntypes,
                    _d(table.eps), _d(table.sigma), _d(table.third), _i(table.types),
                    len(first), _i(first), _i(last), _d(f), _d(fabs), _d(U), _d(W),
                    _d(S), nc.ctypes.data_as(_up), no.ctypes.data_as(_up))
    return dict(f=f, fabs=fabs, U=float(U[0]), W=float(W[0]), S=S,
                contacts=int(nc[0]), overlaps=int(no[0]))


def damped_constants(dt, damping):
    c = np.zeros(3)
    lib().or_damped_constants(float(dt), float(damping), _d(c[0:]), _d(c[1:]),
                              _d(c[2:]))
    return tuple(float(v) for v in c)


class System:
    """Whole-system restatement: AtomVec + OriginBox + NeighborList + one
    NListed interaction + CollectionVerlet / CollectionDamped."""

    def __init__(self, masses, L, table: PairTable, diam, skin, member=None,
                 ndim=3, brute=False):
        masses = _f64(masses).ravel()
        self.n = masses.size
        self.ndim = ndim
        self.table = table
        mem = _members(self.n, member)
        self._keep = (masses, _f64(L).ravel(), mem, _f64(diam).ravel())
        self.h = lib().or_sys_create(ndim, self.n, _d(masses), _d(self._keep[1]),
                                     table.kind, table.ntypes, _d(table.eps),
                                     _d(table.sigma), _d(table.third),
                                     _i(table.types), len(mem), _i(mem),
                                     _d(self._keep[3]), float(skin), int(brute))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_sys_destroy(self.h)
            self.h = None

    def set_state(self, x=None, v=None, a=None, f=None):
        sh = (self.n, self.ndim)
        arrs = [None if q is None else _f64(q, sh) for q in (x, v, a, f)]
        lib().or_sys_set_state(self.h, *[None if q is None else _d(q) for q in arrs])

    def state(self):
        sh = (self.n, self.ndim)
        x, v, a, f = (np.zeros(sh) for _ in range(4))
        lib().or_sys_get_state(self.h, _d(x), _d(v), _d(a), _d(f))
        return x, v, a, f

    def ignore(self, a, b):
        lib().or_sys_ignore(self.h, int(a), int(b))

    def ignore_size(self):
        return int(lib().or_sys_ignore_size(self.h))

    def update_list(self, force=True):
        return bool(lib().or_sys_update_list(self.h, int(force)))

    def which(self):
        return int(lib().or_sys_which(self.h))

    def numpairs(self):
        return int(lib().or_sys_numpairs(self.h))

    def pairs(self):
        n = self.numpairs()
        first = np.empty(max(n, 1), np.int32)
        last = np.empty(max(n, 1), np.int32)
        lib().or_sys_get_pairs(self.h, n, _i(first), _i(last))
        return first[:n], last[:n]

    def set_forces(self, and_a=True):
        lib().or_sys_set_forces(self.h, int(and_a))

    def potential_energy(self):
        return float(lib().or_sys_potential_energy(self.h))

    def virial(self):
        return float(lib().or_sys_virial(self.h))

    def stress(self):
        S = np.zeros((self.ndim, self.ndim))
        lib().or_sys_stress(self.h, _d(S))
        return S

    def kinetic_energy(self, originv=None):
        ov = None if originv is None else _d(_f64(originv))
        return float(lib().or_sys_kinetic_energy(self.h, ov))

    def momentum(self):
        p = np.zeros(self.ndim)
        lib().or_sys_momentum(self.h, _d(p))
        return p

    def mass(self):
        return float(lib().or_sys_mass(self.h))

    def temp(self, minuscomv=True):
        return float(lib().or_sys_temp(self.h, int(minuscomv)))

    def energy(self):
        return self.potential_energy() + self.kinetic_energy()

    def initialize(self):
        lib().or_sys_initialize(self.h)

    def verlet(self, dt, nsteps=1):
        lib().or_sys_verlet(self.h, float(dt), int(nsteps))

    def damped(self, dt, damping, nsteps=1):
        lib().or_sys_damped(self.h, float(dt), float(damping), int(nsteps))

    def box(self):
        L = np.zeros(self.ndim)
        lib().or_sys_box(self.h, _d(L))
        return L


class NLCG:
    """CollectionNLCG (or CollectionNLCGFixedL) on an oracle System; creating
    it runs Collection::initialize on the system."""

    INFO = ("alpha", "beta", "betaused", "dxsum", "alphavmax", "sec", "k", "vl", "fl",
            "al", "Knew", "maxdV")

    def __init__(self, system, dt, P0, kappa=10.0, kmax=1000, secmax=40, seceps=1e-20,
                 fixed_L=False):
        self.system = system
        self.h = lib().or_nlcg_create(system.h, float(dt), float(P0), float(kappa),
                                      float(kmax), int(secmax), float(seceps),
                                      int(bool(fixed_L)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_nlcg_destroy(self.h)
            self.h = None

    def set_limits(self, alphamax=2.0, afrac=0.0, dxmax=100.0, stepmax=1e-3, maxdV=0.0):
        lib().or_nlcg_set_limits(self.h, float(alphamax), float(afrac), float(dxmax),
                                 float(stepmax), float(maxdV))

    def set_forces(self, constraints_and_a=True, setV=True):
        lib().or_nlcg_set_forces(self.system.h, self.h, int(bool(constraints_and_a)),
                                 int(bool(setV)))

    def reset(self):
        lib().or_nlcg_reset(self.system.h, self.h)

    def timestep(self, nsteps=1):
        lib().or_nlcg_timestep(self.system.h, self.h, int(nsteps))

    def descend(self):
        lib().or_nlcg_descend(self.system.h, self.h)

    def observables(self):
        H, P, K = np.zeros(1), np.zeros(1), np.zeros(1)
        lib().or_nlcg_observables(self.system.h, self.h, _d(H), _d(P), _d(K))
        return dict(H=float(H[0]), P=float(P[0]), K=float(K[0]))

    def info(self):
        out = np.zeros(len(self.INFO))
        lib().or_nlcg_info(self.h, _d(out))
        return dict(zip(self.INFO, out.tolist()))
