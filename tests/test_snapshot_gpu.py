"""Asynchronous state snapshots and xyz frames (parm_snapshot_*): the copy
leaves the device while later steps run; what it holds is the state at the
time it was taken; the xyz frame is the reference drivers' format
(src/bin/LJatoms.cpp:130-158: count, comment, then element and
diff(0, x) = remainder(-x, L) per atom, tab-separated, ostream %g)."""
import math

import numpy as np
import pytest

from parm_amd import capi
from tests import systems as S

pytestmark = pytest.mark.gpu


def _device_or_skip():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible: the -m gpu tests need an MI355X")


def _lj(n=3000, ndim=3):
    x, L = S.jittered_lattice(n, 0.8, ndim, seed=n)
    p = S.lj_params(n)
    s = capi.System(ndim, np.ones(n))
    s.set_box(L)
    s.nlist_set(0.3, p["diam"])
    s.set_lj(p["eps"], p["sigma"], p["third"], p["types"])
    s.set_state(x=x, v=S.gaussian_velocities(n, ndim, 1.0, seed=2), a=np.zeros_like(x))
    s.initialize()
    return s


def test_snapshot_holds_the_state_when_taken():
    _device_or_skip()
    s = _lj()
    x0, v0, _, _ = s.state()
    t0 = s.snapshot(True, True)
    s.verlet(0.002, 20)          # queued behind the gather; the copy overlaps
    t1 = s.snapshot(True, False)
    x1 = s.positions()
    xs, vs, L = s.snapshot_read(t0, True, True)
    assert np.array_equal(xs, x0) and np.array_equal(vs, v0)
    assert np.array_equal(L, s.box())
    xs1, _, _ = s.snapshot_read(t1)
    assert np.array_equal(xs1, x1) and not np.array_equal(xs1, x0)
    with pytest.raises(capi.ParmError):
        s.snapshot_read(t1, velocities=True)
    # only the last four are held
    for _ in range(4):
        s.snapshot()
    with pytest.raises(ValueError):
        s.snapshot_read(t0)
    s.close()


def test_xyz_frames(tmp_path):
    _device_or_skip()
    s = _lj(n=500)
    path = tmp_path / "traj.xyz"
    frames = []
    for q in range(3):
        t = s.snapshot()
        x, _, L = s.snapshot_read(t)
        frames.append((x, L))
        s.write_xyz(t, str(path), append=q > 0, comment=f"frame {q}", precision=6)
        s.verlet(0.002, 5)
    t = s.snapshot()
    s.write_xyz(t, str(tmp_path / "one.xyz"), append=False, comment=None,
                elements="CO" * 250, precision=24)
    lines = path.read_text().split("\n")
    assert lines[-1] == ""
    per = 2 + 500
    assert len(lines) - 1 == 3 * per
    for q, (x, L) in enumerate(frames):
        blk = lines[q * per:(q + 1) * per]
        assert blk[0] == "500" and blk[1] == f"frame {q}"
        for i in (0, 17, 499):
            parts = blk[2 + i].split("\t")
            assert parts[0] == "C" and len(parts) == 4
            for k in range(3):
                assert parts[1 + k] == "%.6g" % math.remainder(-x[i, k], L[k])
    one = (tmp_path / "one.xyz").read_text().split("\n")
    assert one[1] == "" and one[2].startswith("C\t") and one[3].startswith("O\t")
    xo, _, Lo = s.snapshot_read(t)
    vals = [float(c) for c in one[3].split("\t")[1:]]
    assert vals == [math.remainder(-xo[1, k], Lo[k]) for k in range(3)]
    s.close()
