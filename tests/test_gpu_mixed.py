"""Mixed-precision walk (FP32 displacements and interactions, FP64 Kahan
sums across batches, FP64 state) against the plain FP64 path on the same
tree and against the FP64 direct sum.  Tolerances: median relative
acceleration error <= 1e-6 and 99th percentile <= 1e-5 versus the FP64 tree
(the accuracy target for the mixed walk is 1e-5 median against the direct sum)."""
import numpy as np
import pytest

from helpers import rel_err
from milkywayathome_client_amd import dwarf, models, nbody as nb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dwarf_case():
    pot = nb.milky_way_potential()
    x0 = models.lbr_to_cartesian([218.0, 53.5, 28.6])
    fpos, fvel = models.reverse_orbit(pot, x0, [-156.0, 79.0, 107.0], 4.0, 1e-4)
    pos, vel, m, _ = dwarf.two_component_dwarf(40000, seed=34086709, position=fpos,
                                               velocity=fvel)
    eps2 = models.calculate_eps2(40000, 0.2, 0.8, 12.0, 60.0)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=eps2, theta=0.5)
    return ctx, pot, pos, vel, m


def test_mixed_matches_fp64_tree(dwarf_case):
    ctx, pot, pos, vel, m = dwarf_case
    a64 = nb.NBodyState(ctx, pot, pos, vel, m).accelerations()
    a32 = nb.NBodyState(ctx, pot, pos, vel, m, precision="mixed").accelerations()
    e = rel_err(a32, a64, 0)
    print("total: median", np.median(e), "p99", np.percentile(e, 99))
    assert np.median(e) <= 1e-6 and np.percentile(e, 99) <= 1e-5
    s64 = nb.NBodyState(ctx, None, pos, vel, m)
    s32 = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed")
    e = rel_err(s32.accelerations(), s64.accelerations(), 0)
    print("self-gravity: median", np.median(e), "p99", np.percentile(e, 99))
    assert np.median(e) <= 1e-6 and np.percentile(e, 99) <= 1e-5


def test_mixed_error_vs_direct_sum_is_the_tree_error(dwarf_case):
    ctx, pot, pos, vel, m = dwarf_case
    idx = np.arange(0, pos.shape[0], 10, dtype=np.int32)
    errs = {}
    for prec in ("fp64", "mixed"):
        st = nb.NBodyState(ctx, None, pos, vel, m, precision=prec)
        d = st.direct_accel(idx)
        errs[prec] = np.median(rel_err(st.accelerations()[idx], d, 0))
    print(errs)
    assert abs(errs["mixed"] - errs["fp64"]) <= 1e-6


def test_mixed_interaction_lists_match_fp64(dwarf_case):
    ctx, pot, pos, vel, m = dwarf_case
    c64 = nb.NBodyState(ctx, pot, pos, vel, m).walk_counters()
    c32 = nb.NBodyState(ctx, pot, pos, vel, m, precision="mixed").walk_counters()
    for a, b in zip(c64[:3], c32[:3]):
        assert abs(a - b) <= 1e-5 * a          # FP32 opening tests flip rarely


def test_mixed_trajectories_track_fp64(dwarf_case):
    ctx, pot, pos, vel, m = dwarf_case
    s64 = nb.NBodyState(ctx, pot, pos, vel, m)
    s32 = nb.NBodyState(ctx, pot, pos, vel, m, precision="mixed")
    s64.steps(50)
    s32.steps(50)
    p64, v64, _ = s64.bodies()
    p32, v32, _ = s32.bodies()
    dx = np.linalg.norm(p32 - p64, axis=1)
    print("50 steps: median |dx|", np.median(dx), "max", dx.max())
    assert np.median(dx) < 1e-6


def test_mixed_handles_massless_and_tiny():
    pos = np.array([[0.0, 0.0, 0.0], [0.5, 0.1, 0.0], [1.0, 1.0, 1.0]])
    vel = np.zeros_like(pos)
    m = np.array([1.0, 1.0, 0.0])
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-4, theta=0.5,
                             allowIncest=True)
    a64 = nb.NBodyState(ctx, None, pos, vel, m).accelerations()
    a32 = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed").accelerations()
    np.testing.assert_allclose(a32, a64, rtol=1e-6, atol=1e-9)


def test_coherent_mixed_is_at_least_as_accurate(dwarf_case):
    """Wave-coherent opening refines every body's list: its error against
    the FP64 direct sum is no larger than the FP64 tree's, it does more
    interactions (each accepted node passes the body's own test, so there
    are no fewer), and it does not depend on the workgroup size."""
    ctx, pot, pos, vel, m = dwarf_case
    idx = np.arange(0, pos.shape[0], 10, dtype=np.int32)
    st64 = nb.NBodyState(ctx, None, pos, vel, m)
    d = st64.direct_accel(idx)
    e64 = np.median(rel_err(st64.accelerations()[idx], d, 0))
    coh = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed_coherent")
    ecoh = np.median(rel_err(coh.accelerations()[idx], d, 0))
    print("median vs direct: fp64 tree", e64, "coherent mixed", ecoh)
    assert ecoh <= e64
    c64, ccoh = st64.walk_counters(), coh.walk_counters()
    assert ccoh[0] + ccoh[1] >= c64[0] + c64[1]
    coh.set_tune(64, 1024)
    a = coh.walk_accel()
    coh.set_tune(256, 1024)
    assert np.array_equal(coh.walk_accel(), a)


@pytest.mark.parametrize("criterion,theta", [("sw93", 0.6), ("treecode", 0.5), ("bh86", 0.0)])
def test_mixed_other_criteria_match_fp64(dwarf_case, criterion, theta):
    """The mixed walk reads each cell's critical radius from its record, so
    every opening criterion (SW93, TreeCode, theta = 0) follows the FP64
    walk's lists and forces to the same tolerance as BH86."""
    _, pot, pos, vel, m = dwarf_case
    n = 20000
    pos, vel, m = pos[:n], vel[:n], m[:n]
    eps2 = models.calculate_eps2(n, 0.2, 0.8, 12.0, 60.0)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=eps2, theta=theta,
                             criterion=criterion)
    s64 = nb.NBodyState(ctx, pot, pos, vel, m)
    s32 = nb.NBodyState(ctx, pot, pos, vel, m, precision="mixed")
    e = rel_err(s32.accelerations(), s64.accelerations(), 0)
    print(criterion, theta, "median", np.median(e), "p99", np.percentile(e, 99))
    assert np.median(e) <= 1e-6 and np.percentile(e, 99) <= 1e-5
    c64, c32 = s64.walk_counters(), s32.walk_counters()
    for a, b in zip(c64[:3], c32[:3]):
        assert abs(a - b) <= 1e-4 * max(a, 1)
    s64.close()
    s32.close()


def test_coherent_open_scale_refines(dwarf_case):
    """nbhip_set_open_scale: scaling the critical radii of the coherent walk
    opens more cells, so it does more interactions and its error against the
    direct sum falls; scale 1 restores the default forces bit for bit; other
    walks and scales below 1 are rejected."""
    ctx, pot, pos, vel, m = dwarf_case
    idx = np.arange(0, pos.shape[0], 10, dtype=np.int32)
    coh = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed_coherent")
    d = coh.direct_accel(idx)
    a1 = coh.accelerations()
    e1 = np.median(rel_err(a1[idx], d, 0))
    c1 = coh.walk_counters()
    coh.set_open_scale(2.0)
    coh.gravity()
    e2 = np.median(rel_err(coh.accelerations()[idx], d, 0))
    c2 = coh.walk_counters()
    print("median vs direct: scale 1", e1, "scale 2", e2, "interactions", c1[:2], c2[:2])
    assert e2 < e1 and c2[0] + c2[1] > c1[0] + c1[1]
    coh.set_open_scale(1.0)
    coh.gravity()
    assert np.array_equal(coh.accelerations(), a1)
    with pytest.raises(nb.NBodyError):
        coh.set_open_scale(0.5)
    coh.close()
    mixed = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed", prestep=False)
    with pytest.raises(nb.NBodyError):
        mixed.set_open_scale(2.0)
    mixed.close()
