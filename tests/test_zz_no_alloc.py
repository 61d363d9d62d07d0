"""nbhip_histogram_noalloc, the allocation-free variant of nbhip_histogram,
called through the C ABI: the same counts and totals as nbhip_histogram on
two calls in a row, the second with a smaller input (fewer simulations,
bodies and bins), so counters left from the first call must not leak into
the second; sizes beyond its fixed scratch are rejected."""
import ctypes as C

import numpy as np
import pytest

from helpers import plummer_bodies
from milkywayathome_client_amd import histogram as H, models, nbody as nb

pytestmark = pytest.mark.gpu


def _hp(lambda_bins, beta_bins):
    return H.HistogramParams(lambdaBins=lambda_bins, betaBins=beta_bins)


def _c_hp(hp):
    return nb._HistParams(hp.phi, hp.theta, hp.psi, hp.lambdaStart, hp.lambdaEnd, hp.betaStart,
                          hp.betaEnd, hp.lambdaBins, hp.betaBins, hp.sunGCDist)


def _noalloc(st, hp):
    nsim = getattr(st, "nsim", 1)
    counts = np.full((nsim, hp.lambdaBins * hp.betaBins), 0xDEADBEEF, dtype=np.uint32)
    totals = np.full(nsim, 0xDEADBEEF, dtype=np.uint32)
    h = _c_hp(hp)
    rc = st.lib.nbhip_histogram_noalloc(st._h, C.byref(h),
                                        counts.ctypes.data_as(C.POINTER(C.c_uint)),
                                        totals.ctypes.data_as(C.POINTER(C.c_uint)))
    return rc, counts, totals


def _member(n, seed, radius):
    # spread along the orbit's neighbourhood so that many lambda bins fill
    centre = models.lbr_to_cartesian([218.0, 53.5, 28.6])
    return plummer_bodies(n, seed=seed, mass=10.0, radius=radius, position=tuple(centre),
                          velocity=(-156.0, 79.0, 107.0))


def test_noalloc_matches_histogram_on_two_calls():
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-4, theta=0.5)
    pot = nb.milky_way_potential()
    # first call: three simulations, 50 x 4 bins, the second half of each ignored
    members = [_member(n, 30 + k, 3.0 + k) for k, n in enumerate((6000, 9000, 4000))]
    bt = nb.NBodyBatch(ctx, pot, members)
    bt.set_ignore(np.concatenate([np.arange(len(mm[2])) >= len(mm[2]) // 2 for mm in members]))
    hp1 = _hp(50, 4)
    print("call 1: batch of 3, 50 x 4 bins", flush=True)
    rc, c1, t1 = _noalloc(bt, hp1)
    assert rc == 0, bt.lib.nbhip_last_error()
    r1, rt1 = bt.histogram(hp1)
    assert np.array_equal(c1, r1) and np.array_equal(t1, rt1)
    assert t1.min() > 100 and (c1 > 0).sum() > 20
    # second call: one smaller state, 30 x 1 bins, after a few steps
    pos, vel, m = _member(2500, 77, 2.0)
    st = nb.NBodyState(ctx, pot, pos, vel, m)
    assert st.steps(3) == 0
    hp2 = _hp(30, 1)
    print("call 2: one simulation, 30 x 1 bins", flush=True)
    rc, c2, t2 = _noalloc(st, hp2)
    assert rc == 0, st.lib.nbhip_last_error()
    r2, rt2 = st.histogram(hp2)
    assert np.array_equal(c2, r2) and np.array_equal(t2, rt2)
    assert int(t2[0]) == int(c2.sum()) and t2[0] > 100
    # and the first state once more: the second call's counters are cleared too
    rc, c3, t3 = _noalloc(bt, hp1)
    assert rc == 0 and np.array_equal(c3, r1) and np.array_equal(t3, rt1)
    st.close()
    bt.close()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-4, theta=0.5)
    members = [_member(500, 40 + k, 2.0) for k in range(3)]
    bt = nb.NBodyBatch(ctx, None, members)
    # 3 x (1000 x 400 + 1) counters > 2^20
    hp = _hp(1000, 400)
    rc, _, _ = _noalloc(bt, hp)
    assert rc == nb.NBODY_USER_ERROR
    assert b"exceed the scratch" in bt.lib.nbhip_last_error()
    # the largest size that fits still works and matches
    hp = _hp(349525, 1)            # 3 x 349526 = 1048578 > 2^20: one bin fewer fits
    rc, _, _ = _noalloc(bt, hp)
    assert rc == nb.NBODY_USER_ERROR
    hp = _hp(349524, 1)
    rc, c, t = _noalloc(bt, hp)
    assert rc == 0
    r, rt = bt.histogram(hp)
    assert np.array_equal(c, r) and np.array_equal(t, rt)
    bt.close()
