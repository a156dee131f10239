"""parm_pair_sums_noalloc (the entry point that allocates nothing) against
parm_pair_sums, through the C ABI: the same U, W, S, contacts and overlaps on
two calls in a row, the second on a smaller system (the fixed scratch keeps
the first call's contents, which must not leak into the second)."""
import ctypes as C

import numpy as np
import pytest

from parm_amd import capi
from tests import systems as S

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def _system(n, seed, kind):
    x, L = S.jittered_lattice(n, 0.9, 3, seed=seed, jitter=0.1)
    s = capi.System(3, np.ones(n))
    s.set_box(L)
    if kind == "lj":
        p = S.lj_params(n)
        s.nlist_set(0.3, p["diam"])
        s.set_lj(p["eps"], p["sigma"], p["third"], p["types"])
    else:
        p = S.bidisperse_repulsion(n, seed)
        s.nlist_set(0.3, p["diam"])
        s.set_repulsion(p["eps"], p["sigma"], p["third"], p["types"])
    s.set_state(x=x)
    s.update_list(True)
    return s


def test_noalloc_sums_match_pair_sums_twice():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible")
    hip = _hip()
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), 13 * 8) == 0
    try:
        for n, seed, kind in ((20000, 1, "lj"), (3000, 2, "rep")):
            print(f"noalloc case n={n} kind={kind}", flush=True)
            s = _system(n, seed, kind)
            ref = s.pair_sums()
            capi._check(capi.lib().parm_pair_sums_noalloc(s.h, dev))
            s.synchronize()
            got = np.zeros(13)
            assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), dev, 13 * 8, 2) == 0
            assert got[0] == ref["U"] and got[1] == ref["W"]
            assert np.array_equal(got[2:11], ref["S"].ravel())
            assert got[11] == ref["contacts"] and got[12] == ref["overlaps"]
            s.close()
    finally:
        hip.hipFree(dev)


def test_noalloc_refuses_oversized_systems():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible")
    hip = _hip()
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), 13 * 8) == 0
    try:
        n = (1 << 21) + 1
        s = capi.System(3, np.ones(n))
        with pytest.raises(capi.ParmError, match="at most"):
            capi._check(capi.lib().parm_pair_sums_noalloc(s.h, dev))
        s.close()
    finally:
        hip.hipFree(dev)
