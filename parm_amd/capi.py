"""ctypes binding of libparm_hip.so (the C ABI in include/parm_hip.h).

Loads the in-tree library built by ``__graft_entry__.build()``.  There is no
fallback: if the library is missing, or no gfx950 device is present, the
calls raise.  Errors map to the exceptions the reference's SWIG layer raises
for the same C++ exceptions (src/sim.i:139-157): invalid arguments ->
ValueError, everything else -> RuntimeError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libparm_hip.so")

OK, E_INVALID, E_DEVICE, E_NOMEM, E_UNSUPPORTED, E_NONFINITE, E_STATE = 0, -1, -2, -3, -4, -5, -6
FP64, MIXED = 0, 1
PAIR_LJ_ATTRACT_REPULSE, PAIR_REPULSION = 1, 2

#: every function include/parm_hip.h declares (checked by the CPU tests)
EXPORTS = (
    "parm_last_error", "parm_version", "parm_device_count",
    "parm_system_create", "parm_system_destroy", "parm_system_info",
    "parm_set_box", "parm_set_state", "parm_get_state", "parm_set_masses",
    "parm_nlist_set", "parm_nlist_update", "parm_nlist_which",
    "parm_nlist_numpairs", "parm_nlist_pairs", "parm_nlist_ignore",
    "parm_nlist_ignore_size",
    "parm_pair_lj_attract_repulse", "parm_pair_repulsion", "parm_pair_lj_cut",
    "parm_pair_lj_cut_indexed", "parm_pair_lj_repulse", "parm_pair_repulsion_indexed",
    "parm_set_forces", "parm_add_pair_forces", "parm_pair_sums", "parm_kinetic_energy", "parm_momentum",
    "parm_force_stats", "parm_initialize", "parm_verlet_steps",
    "parm_damped_steps", "parm_synchronize", "parm_profile_enable",
    "parm_profile_read", "parm_list_info", "parm_autotune", "parm_get_box",
    "parm_nlcg_create", "parm_nlcg_destroy", "parm_nlcg_set_limits", "parm_nlcg_set_params",
    "parm_nlcg_reset", "parm_nlcg_set_forces", "parm_nlcg_timestep", "parm_nlcg_descend",
    "parm_nlcg_observables", "parm_nlcg_info",
    "parm_snapshot_begin", "parm_snapshot_read", "parm_snapshot_write_xyz",
    "parm_pair_sums_noalloc",
    "parm_batch_create", "parm_batch_destroy", "parm_batch_info",
    "parm_batch_set_masses", "parm_batch_set_state", "parm_batch_get_state",
    "parm_batch_pair_repulsion", "parm_batch_pair_lj_attract_repulse",
    "parm_batch_initialize", "parm_batch_damped_steps", "parm_batch_verlet_steps",
    "parm_batch_sums", "parm_batch_lists", "parm_batch_pairs",
    "parm_batch_pair_sums", "parm_batch_autotune", "parm_batch_profile_enable",
    "parm_batch_profile_read",
)


class ParmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (parm status {code})")
        self.code = code


class ParmValueError(ValueError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (parm status {code})")
        self.code = code


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p


def lib():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(the device path has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        i64, i32, dbl = C.c_int64, C.c_int, C.c_double
        sig = {
            "parm_last_error": (C.c_char_p, []),
            "parm_version": (i32, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "parm_device_count": (i32, [C.POINTER(C.c_int)]),
            "parm_system_create": (i32, [i32, i32, i64, _dp, C.POINTER(_vp)]),
            "parm_system_destroy": (i32, [_vp]),
            "parm_system_info": (i32, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int64)]),
            "parm_set_box": (i32, [_vp, _dp]),
            "parm_set_state": (i32, [_vp, _dp, _dp, _dp, _dp]),
            "parm_get_state": (i32, [_vp, _dp, _dp, _dp, _dp]),
            "parm_set_masses": (i32, [_vp, _dp]),
            "parm_nlist_set": (i32, [_vp, dbl, i64, _ip, _dp]),
            "parm_nlist_update": (i32, [_vp, i32, C.POINTER(C.c_int)]),
            "parm_nlist_which": (i32, [_vp, C.POINTER(C.c_uint32)]),
            "parm_nlist_numpairs": (i32, [_vp, C.POINTER(C.c_int64)]),
            "parm_nlist_pairs": (i32, [_vp, i64, _ip, _ip, C.POINTER(C.c_int64)]),
            "parm_nlist_ignore": (i32, [_vp, i64, _ip, _ip]),
            "parm_nlist_ignore_size": (i32, [_vp, C.POINTER(C.c_int64)]),
            "parm_pair_lj_attract_repulse": (i32, [_vp, i32, _dp, _dp, _dp, _ip]),
            "parm_pair_repulsion": (i32, [_vp, i32, _dp, _dp, _dp, _ip]),
            "parm_pair_lj_cut": (i32, [_vp, i32, _dp, _dp, _dp, _ip]),
            "parm_pair_lj_cut_indexed": (i32, [_vp, i32, _dp, _dp, _dp, _ip]),
            "parm_pair_lj_repulse": (i32, [_vp, i32, _dp, _dp, _ip]),
            "parm_pair_repulsion_indexed": (i32, [_vp, i32, _dp, _dp, _dp, _ip]),
            "parm_set_forces": (i32, [_vp, i32]),
            "parm_add_pair_forces": (i32, [_vp]),
            "parm_pair_sums": (i32, [_vp, _dp, _dp, _dp, C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint64)]),
            "parm_kinetic_energy": (i32, [_vp, _dp, _dp]),
            "parm_momentum": (i32, [_vp, _dp, _dp]),
            "parm_force_stats": (i32, [_vp, _dp, _dp]),
            "parm_initialize": (i32, [_vp]),
            "parm_verlet_steps": (i32, [_vp, dbl, i64]),
            "parm_damped_steps": (i32, [_vp, dbl, dbl, i64]),
            "parm_synchronize": (i32, [_vp]),
            "parm_profile_enable": (i32, [_vp, i32]),
            "parm_profile_read": (i32, [_vp, i32, _dp, C.POINTER(C.c_int64), _dp,
                                        C.POINTER(C.c_int64)]),
            "parm_list_info": (i32, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]),
            "parm_autotune": (i32, [_vp, dbl, i32]),
            "parm_get_box": (i32, [_vp, _dp]),
            "parm_nlcg_create": (i32, [_vp, dbl, dbl, dbl, dbl, i32, dbl, i32,
                                       C.POINTER(_vp)]),
            "parm_nlcg_destroy": (i32, [_vp]),
            "parm_nlcg_set_limits": (i32, [_vp, dbl, dbl, dbl, dbl, dbl]),
            "parm_nlcg_set_params": (i32, [_vp, dbl, dbl, dbl, dbl, i32, dbl]),
            "parm_nlcg_reset": (i32, [_vp]),
            "parm_nlcg_set_forces": (i32, [_vp, i32, i32]),
            "parm_nlcg_timestep": (i32, [_vp, i64]),
            "parm_nlcg_descend": (i32, [_vp]),
            "parm_nlcg_observables": (i32, [_vp, _dp, _dp, _dp]),
            "parm_nlcg_info": (i32, [_vp, _dp]),
            "parm_snapshot_begin": (i32, [_vp, i32, C.POINTER(C.c_int64)]),
            "parm_pair_sums_noalloc": (i32, [_vp, _vp]),
            "parm_snapshot_read": (i32, [_vp, i64, _dp, _dp, _dp]),
            "parm_snapshot_write_xyz": (i32, [_vp, i64, C.c_char_p, i32, C.c_char_p,
                                              C.c_char_p, i32]),
            "parm_batch_create": (i32, [i32, i32, i64, C.POINTER(C.c_int64), _dp, dbl,
                                        C.POINTER(_vp)]),
            "parm_batch_destroy": (i32, [_vp]),
            "parm_batch_info": (i32, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int64)]),
            "parm_batch_set_masses": (i32, [_vp, _dp]),
            "parm_batch_set_state": (i32, [_vp, _dp, _dp, _dp]),
            "parm_batch_get_state": (i32, [_vp, _dp, _dp, _dp, _dp]),
            "parm_batch_pair_repulsion": (i32, [_vp, i32, _dp, _dp, _dp, _ip]),
            "parm_batch_pair_lj_attract_repulse": (i32, [_vp, i32, _dp, _dp, _dp, _ip]),
            "parm_batch_initialize": (i32, [_vp]),
            "parm_batch_damped_steps": (i32, [_vp, dbl, dbl, i64, i32]),
            "parm_batch_verlet_steps": (i32, [_vp, dbl, i64, i32]),
            "parm_batch_sums": (i32, [_vp, _dp, _dp]),
            "parm_batch_lists": (i32, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_int64)]),
            "parm_batch_pairs": (i32, [_vp, i64, i64, _ip, _ip, C.POINTER(C.c_int64)]),
            "parm_batch_pair_sums": (i32, [_vp, _dp, _dp, _dp, _dp]),
            "parm_batch_autotune": (i32, [_vp]),
            "parm_batch_profile_enable": (i32, [_vp, i32]),
            "parm_batch_profile_read": (i32, [_vp, i32, _dp, C.POINTER(C.c_int64), _dp,
                                              C.POINTER(C.c_int64)]),
        }
        for name, (res, args) in sig.items():
            # a library older than this binding lacks the newest entry points:
            # those fail when called (AttributeError), the rest stay usable
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        msg = lib().parm_last_error().decode(errors="replace")
        if rc == E_INVALID:
            raise ParmValueError(rc, msg)
        raise ParmError(rc, msg)


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().parm_device_count(C.byref(n)))
    return n.value


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


class System:
    """One device system: AtomVec + OriginBox + NeighborList + one NListed
    pair interaction (see include/parm_hip.h)."""

    def __init__(self, ndim, masses, precision=FP64):
        masses = _f64(masses).ravel()
        self.ndim = int(ndim)
        self.n = masses.size
        self.precision = precision
        h = _vp()
        _check(lib().parm_system_create(self.ndim, int(precision), self.n, _d(masses),
                                        C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().parm_system_destroy(self.h)
            self.h = None

    __del__ = close

    # -- state
    def set_box(self, L):
        L = _f64(np.broadcast_to(np.asarray(L, float), (self.ndim,)))
        _check(lib().parm_set_box(self.h, _d(L)))

    def set_state(self, x=None, v=None, a=None, f=None):
        sh = (self.n, self.ndim)
        arrs = [None if q is None else _f64(q, sh) for q in (x, v, a, f)]
        _check(lib().parm_set_state(self.h, *[_d(q) for q in arrs]))

    def state(self):
        sh = (self.n, self.ndim)
        x, v, a, f = (np.empty(sh) for _ in range(4))
        _check(lib().parm_get_state(self.h, _d(x), _d(v), _d(a), _d(f)))
        return x, v, a, f

    def positions(self):
        x = np.empty((self.n, self.ndim))
        _check(lib().parm_get_state(self.h, _d(x), None, None, None))
        return x

    def forces(self):
        f = np.empty((self.n, self.ndim))
        _check(lib().parm_get_state(self.h, None, None, None, _d(f)))
        return f

    def set_masses(self, m):
        _check(lib().parm_set_masses(self.h, _d(_f64(m).ravel())))

    # -- neighbour list
    def nlist_set(self, skin, diameters, members=None):
        diam = _f64(diameters).ravel()
        mem = np.arange(diam.size, dtype=np.int32) if members is None else \
            np.ascontiguousarray(members, dtype=np.int32)
        _check(lib().parm_nlist_set(self.h, float(skin), mem.size,
                                    mem.ctypes.data_as(_ip), _d(diam)))

    def update_list(self, force=True):
        r = C.c_int(0)
        _check(lib().parm_nlist_update(self.h, int(bool(force)), C.byref(r)))
        return bool(r.value)

    def ignore(self, a, b):
        """NeighborList::ignore for each (a[k], b[k]) pair of atom indices."""
        a = np.atleast_1d(np.ascontiguousarray(a, dtype=np.int32))
        b = np.atleast_1d(np.ascontiguousarray(b, dtype=np.int32))
        _check(lib().parm_nlist_ignore(self.h, a.size, a.ctypes.data_as(_ip),
                                       b.ctypes.data_as(_ip)))

    def ignore_size(self):
        n = C.c_int64(0)
        _check(lib().parm_nlist_ignore_size(self.h, C.byref(n)))
        return n.value

    def which(self):
        w = C.c_uint32(0)
        _check(lib().parm_nlist_which(self.h, C.byref(w)))
        return w.value

    def numpairs(self):
        n = C.c_int64(0)
        _check(lib().parm_nlist_numpairs(self.h, C.byref(n)))
        return n.value

    def pairs(self):
        n = self.numpairs()
        first = np.empty(max(n, 1), np.int32)
        last = np.empty(max(n, 1), np.int32)
        got = C.c_int64(0)
        _check(lib().parm_nlist_pairs(self.h, n, first.ctypes.data_as(_ip),
                                      last.ctypes.data_as(_ip), C.byref(got)))
        return first[:n], last[:n]

    # -- interactions
    def set_lj(self, eps_rows, sigma, sigcut, types):
        sigma = _f64(sigma).ravel()
        nt = sigma.size
        t = np.ascontiguousarray(types, dtype=np.int32)
        _check(lib().parm_pair_lj_attract_repulse(
            self.h, nt, _d(_f64(eps_rows, (nt, nt))), _d(sigma), _d(_f64(sigcut).ravel()),
            t.ctypes.data_as(_ip)))

    def set_repulsion(self, eps, sigma, exponent, types):
        sigma = _f64(sigma).ravel()
        t = np.ascontiguousarray(types, dtype=np.int32)
        _check(lib().parm_pair_repulsion(self.h, sigma.size, _d(_f64(eps).ravel()),
                                         _d(sigma), _d(_f64(exponent).ravel()),
                                         t.ctypes.data_as(_ip)))

    def set_lj_cut(self, eps, sigma, sigcut, types):
        """NListed<EpsSigCutAtom, LennardJonesCutPair>, per-type parameters."""
        sigma = _f64(sigma).ravel()
        t = np.ascontiguousarray(types, dtype=np.int32)
        _check(lib().parm_pair_lj_cut(self.h, sigma.size, _d(_f64(eps).ravel()), _d(sigma),
                                      _d(_f64(sigcut).ravel()), t.ctypes.data_as(_ip)))

    def set_lj_cut_indexed(self, eps_rows, sigma_rows, sigcut, types):
        """NListed<IEpsISigCutAtom, LennardJonesCutPair> (LJIICut)."""
        sigcut = _f64(sigcut).ravel()
        nt = sigcut.size
        t = np.ascontiguousarray(types, dtype=np.int32)
        _check(lib().parm_pair_lj_cut_indexed(self.h, nt, _d(_f64(eps_rows, (nt, nt))),
                                              _d(_f64(sigma_rows, (nt, nt))), _d(sigcut),
                                              t.ctypes.data_as(_ip)))

    def set_lj_repulse(self, eps, sigma, types):
        """NListed<EpsSigAtom, LJRepulsePair> (LJRepulse)."""
        sigma = _f64(sigma).ravel()
        t = np.ascontiguousarray(types, dtype=np.int32)
        _check(lib().parm_pair_lj_repulse(self.h, sigma.size, _d(_f64(eps).ravel()),
                                          _d(sigma), t.ctypes.data_as(_ip)))

    def set_repulsion_indexed(self, eps_rows, sigma_rows, exponent, types):
        """NListed<IEpsISigExpAtom, RepulsionPair> (RepulsionII)."""
        expo = _f64(exponent).ravel()
        nt = expo.size
        t = np.ascontiguousarray(types, dtype=np.int32)
        _check(lib().parm_pair_repulsion_indexed(self.h, nt, _d(_f64(eps_rows, (nt, nt))),
                                                 _d(_f64(sigma_rows, (nt, nt))), _d(expo),
                                                 t.ctypes.data_as(_ip)))

    def set_forces(self, and_a=True):
        _check(lib().parm_set_forces(self.h, int(bool(and_a))))

    def add_pair_forces(self):
        _check(lib().parm_add_pair_forces(self.h))

    def pair_sums(self):
        U, W = C.c_double(0), C.c_double(0)
        S = np.zeros((self.ndim, self.ndim))
        nc, no = C.c_uint64(0), C.c_uint64(0)
        _check(lib().parm_pair_sums(self.h, C.byref(U), C.byref(W), _d(S), C.byref(nc),
                                    C.byref(no)))
        return dict(U=U.value, W=W.value, S=S, contacts=nc.value, overlaps=no.value)

    def kinetic_energy(self, originv=None):
        K = C.c_double(0)
        ov = None if originv is None else _f64(originv).ravel()
        _check(lib().parm_kinetic_energy(self.h, _d(ov), C.byref(K)))
        return K.value

    def momentum(self):
        p = np.zeros(self.ndim)
        m = C.c_double(0)
        _check(lib().parm_momentum(self.h, _d(p), C.byref(m)))
        return p, m.value

    def force_stats(self):
        fmax, fmean = C.c_double(0), C.c_double(0)
        _check(lib().parm_force_stats(self.h, C.byref(fmax), C.byref(fmean)))
        return fmax.value, fmean.value

    # -- integrators
    def initialize(self):
        _check(lib().parm_initialize(self.h))

    def verlet(self, dt, nsteps=1):
        _check(lib().parm_verlet_steps(self.h, float(dt), int(nsteps)))

    def damped(self, dt, damping, nsteps=1):
        _check(lib().parm_damped_steps(self.h, float(dt), float(damping), int(nsteps)))

    def synchronize(self):
        _check(lib().parm_synchronize(self.h))

    def profile_enable(self, on=True, every=1):
        """Time every `every`-th fused step launch and every rebuild."""
        _check(lib().parm_profile_enable(self.h, int(every) if on else 0))

    def profile_read(self, reset=True):
        sm, rm = C.c_double(0), C.c_double(0)
        sn, rn = C.c_int64(0), C.c_int64(0)
        _check(lib().parm_profile_read(self.h, int(bool(reset)), C.byref(sm), C.byref(sn),
                                       C.byref(rm), C.byref(rn)))
        return dict(step_ms=sm.value, step_launches=sn.value, rebuild_ms=rm.value,
                    rebuilds=rn.value)

    def box(self):
        L = np.empty(self.ndim)
        _check(lib().parm_get_box(self.h, _d(L)))
        return L

    # -- asynchronous output snapshots
    def snapshot(self, positions=True, velocities=False):
        t = C.c_int64(0)
        what = (1 if positions else 0) | (2 if velocities else 0)
        _check(lib().parm_snapshot_begin(self.h, what, C.byref(t)))
        return t.value

    def snapshot_read(self, ticket, positions=True, velocities=False):
        sh = (self.n, self.ndim)
        x = np.empty(sh) if positions else None
        v = np.empty(sh) if velocities else None
        L = np.empty(self.ndim)
        _check(lib().parm_snapshot_read(self.h, int(ticket), _d(x), _d(v), _d(L)))
        return x, v, L

    def write_xyz(self, ticket, path, append=True, comment="", elements=None, precision=6):
        el = None if elements is None else "".join(elements).encode()
        _check(lib().parm_snapshot_write_xyz(self.h, int(ticket), os.fsencode(path),
                                             int(bool(append)),
                                             None if comment is None else comment.encode(),
                                             el, int(precision)))

    def autotune(self, dt, damped=False):
        _check(lib().parm_autotune(self.h, float(dt), int(bool(damped))))

    def list_info(self):
        t, l, c = C.c_int(0), C.c_int(0), C.c_int(0)
        _check(lib().parm_list_info(self.h, C.byref(t), C.byref(l), C.byref(c)))
        return dict(tiled=bool(t.value), lanes=l.value, capacity=c.value)


class NLCG:
    """CollectionNLCG (or CollectionNLCGFixedL) on a System (parm_nlcg_* in
    include/parm_hip.h).  Keeps a reference to the system."""

    INFO = ("alpha", "beta", "betaused", "dxsum", "alphavmax", "sec", "k", "vl", "fl",
            "al", "Knew", "maxdV")

    def __init__(self, system, dt, P0, kappa=10.0, kmax=1000, secmax=40, seceps=1e-20,
                 fixed_L=False):
        self.system = system
        h = _vp()
        _check(lib().parm_nlcg_create(system.h, float(dt), float(P0), float(kappa),
                                      float(kmax), int(secmax), float(seceps),
                                      int(bool(fixed_L)), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().parm_nlcg_destroy(self.h)
            self.h = None

    __del__ = close

    def set_limits(self, alphamax=2.0, afrac=0.0, dxmax=100.0, stepmax=1e-3, maxdV=0.0):
        _check(lib().parm_nlcg_set_limits(self.h, float(alphamax), float(afrac),
                                          float(dxmax), float(stepmax), float(maxdV)))

    def set_params(self, dt, P0, kappa, kmax=1000, secmax=40, seceps=1e-20):
        """Sets the parameter members; like the reference's set_dt /
        set_pressure_goal / set_kappa, follow with reset()."""
        _check(lib().parm_nlcg_set_params(self.h, float(dt), float(P0), float(kappa),
                                          float(kmax), int(secmax), float(seceps)))

    def reset(self):
        _check(lib().parm_nlcg_reset(self.h))

    def set_forces(self, constraints_and_a=True, setV=True):
        _check(lib().parm_nlcg_set_forces(self.h, int(bool(constraints_and_a)),
                                          int(bool(setV))))

    def timestep(self, nsteps=1):
        _check(lib().parm_nlcg_timestep(self.h, int(nsteps)))

    def descend(self):
        _check(lib().parm_nlcg_descend(self.h))

    def observables(self):
        H, P, K = C.c_double(0), C.c_double(0), C.c_double(0)
        _check(lib().parm_nlcg_observables(self.h, C.byref(H), C.byref(P), C.byref(K)))
        return dict(H=H.value, P=P.value, K=K.value)

    def info(self):
        out = np.empty(len(self.INFO))
        _check(lib().parm_nlcg_info(self.h, _d(out)))
        return dict(zip(self.INFO, out.tolist()))


class Batch:
    """Many independent systems stepped together (parm_batch_* in
    include/parm_hip.h).  natoms[s] atoms of system s, concatenated in system
    order; L is nsys x ndim box edges."""

    def __init__(self, ndim, natoms, L, skin, precision=FP64):
        self.ndim = int(ndim)
        self.natoms = np.ascontiguousarray(natoms, dtype=np.int64).ravel()
        self.nsys = self.natoms.size
        self.offsets = np.concatenate([[0], np.cumsum(self.natoms)]).astype(np.int64)
        self.n = int(self.offsets[-1])
        self.precision = precision
        L = _f64(np.broadcast_to(np.asarray(L, float).reshape(self.nsys, -1),
                                 (self.nsys, self.ndim)))
        h = _vp()
        _check(lib().parm_batch_create(self.ndim, int(precision), self.nsys,
                                       self.natoms.ctypes.data_as(C.POINTER(C.c_int64)),
                                       _d(L), float(skin), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().parm_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def set_masses(self, m):
        _check(lib().parm_batch_set_masses(self.h, _d(_f64(m).ravel())))

    def set_state(self, x=None, v=None, a=None):
        sh = (self.n, self.ndim)
        arrs = [None if q is None else _f64(q, sh) for q in (x, v, a)]
        _check(lib().parm_batch_set_state(self.h, *[_d(q) for q in arrs]))

    def state(self):
        sh = (self.n, self.ndim)
        x, v, a, f = (np.empty(sh) for _ in range(4))
        _check(lib().parm_batch_get_state(self.h, _d(x), _d(v), _d(a), _d(f)))
        return x, v, a, f

    def set_repulsion(self, eps, sigma, exponent, types):
        sigma = _f64(sigma).ravel()
        t = np.ascontiguousarray(types, dtype=np.int32)
        _check(lib().parm_batch_pair_repulsion(self.h, sigma.size, _d(_f64(eps).ravel()),
                                               _d(sigma), _d(_f64(exponent).ravel()),
                                               t.ctypes.data_as(_ip)))

    def set_lj(self, eps_rows, sigma, sigcut, types):
        sigma = _f64(sigma).ravel()
        nt = sigma.size
        t = np.ascontiguousarray(types, dtype=np.int32)
        _check(lib().parm_batch_pair_lj_attract_repulse(
            self.h, nt, _d(_f64(eps_rows, (nt, nt))), _d(sigma), _d(_f64(sigcut).ravel()),
            t.ctypes.data_as(_ip)))

    def initialize(self):
        _check(lib().parm_batch_initialize(self.h))

    def damped(self, dt, damping, nsteps=1, graph=True):
        _check(lib().parm_batch_damped_steps(self.h, float(dt), float(damping), int(nsteps),
                                             int(bool(graph))))

    def verlet(self, dt, nsteps=1, graph=True):
        _check(lib().parm_batch_verlet_steps(self.h, float(dt), int(nsteps),
                                             int(bool(graph))))

    def sums(self):
        U = np.empty(self.nsys)
        c = np.empty(self.nsys)
        _check(lib().parm_batch_sums(self.h, _d(U), _d(c)))
        return U, c

    def pair_sums(self):
        """Per-system U, virial W, contacts and overlaps."""
        out = [np.empty(self.nsys) for _ in range(4)]
        _check(lib().parm_batch_pair_sums(self.h, *[_d(q) for q in out]))
        return dict(U=out[0], W=out[1], contacts=out[2], overlaps=out[3])

    def autotune(self):
        _check(lib().parm_batch_autotune(self.h))

    def profile_enable(self, on=True):
        _check(lib().parm_batch_profile_enable(self.h, int(bool(on))))

    def profile_read(self, reset=True):
        fm, rm = C.c_double(0), C.c_double(0)
        fn, rn = C.c_int64(0), C.c_int64(0)
        _check(lib().parm_batch_profile_read(self.h, int(bool(reset)), C.byref(fm),
                                             C.byref(fn), C.byref(rm), C.byref(rn)))
        return dict(force_ms=fm.value, force_launches=fn.value, rebuild_ms=rm.value,
                    rebuild_launches=rn.value)

    def lists(self):
        w = np.empty(self.nsys, np.uint32)
        n = np.empty(self.nsys, np.int64)
        _check(lib().parm_batch_lists(self.h, w.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      n.ctypes.data_as(C.POINTER(C.c_int64))))
        return w, n

    def pairs(self, s):
        got = C.c_int64(0)
        _check(lib().parm_batch_pairs(self.h, int(s), 0, None, None, C.byref(got)))
        n = got.value
        first = np.empty(max(n, 1), np.int32)
        last = np.empty(max(n, 1), np.int32)
        _check(lib().parm_batch_pairs(self.h, int(s), n, first.ctypes.data_as(_ip),
                                      last.ctypes.data_as(_ip), C.byref(got)))
        return first[:n], last[:n]
