"""The LDS-staged group walk (precision mixed_group) against its numpy
restatement (oracle/group_ref.py) on the same FP64 tree: identical
interaction counts (the lists are decided by FP64 tests) and accelerations
within FP32 rounding (median <= 1e-6, 99th percentile <= 1e-5 relative);
against the FP64 direct sum it must be more accurate than the reference
lists; repeated, eager and graph-captured walks are bit-identical."""
import numpy as np
import pytest

from helpers import plummer_bodies, rel_err
from milkywayathome_client_amd import dwarf, models, nbody as nb
from oracle import group_ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dwarf_case():
    pot = nb.milky_way_potential()
    x0 = models.lbr_to_cartesian([218.0, 53.5, 28.6])
    fpos, fvel = models.reverse_orbit(pot, x0, [-156.0, 79.0, 107.0], 4.0, 1e-4)
    n = 30000
    pos, vel, m, _ = dwarf.two_component_dwarf(n, seed=34086709, position=fpos, velocity=fvel)
    eps2 = models.calculate_eps2(n, 0.2, 0.8, 12.0, 60.0)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=eps2, theta=0.5)
    return ctx, pot, pos, vel, m


def _model(st, ctx, scale=1.0):
    return group_ref.GroupWalk(st.tree(), st.tree_info().rsize, ctx.eps2, open_scale=scale,
                               use_quad=ctx.useQuad)


@pytest.mark.parametrize("scale", [1.0, 1.5])
def test_group_walk_matches_its_restatement(dwarf_case, scale):
    ctx, pot, pos, vel, m = dwarf_case
    st = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed_group", prestep=False)
    if scale != 1.0:
        st.set_open_scale(scale)
    st.gravity()
    gw = _model(st, ctx, scale)
    groups = gw.groups()
    counts = st.walk_counters()
    nbb, nbc = gw.counts(groups)
    print(f"scale {scale}: {len(groups)} groups; gpu counts {counts[:2]} model {(nbb, nbc)}",
          flush=True)
    assert (counts[0], counts[1]) == (nbb, nbc)
    acc = st.walk_accel()
    rng = np.random.default_rng(2)
    errs = []
    for gi in rng.choice(len(groups), size=40, replace=False):
        a, b = groups[gi]
        ref = gw.forces(a, b, gw.accepted(a, b))
        errs.append(rel_err(acc[gw.order[a:b]], ref, 0))
    e = np.concatenate(errs)
    print(f"FP32 vs FP64 lists: median {np.median(e):.3e} p99 {np.percentile(e, 99):.3e} "
          f"max {e.max():.3e}", flush=True)
    assert np.median(e) <= 1e-6 and np.percentile(e, 99) <= 1e-5
    st.close()


def test_group_walk_is_more_accurate_than_reference_lists(dwarf_case):
    ctx, pot, pos, vel, m = dwarf_case
    idx = np.arange(0, pos.shape[0], 7, dtype=np.int32)
    med = {}
    for prec in ("fp64", "mixed_group"):
        st = nb.NBodyState(ctx, pot, pos, vel, m, precision=prec)
        d = st.direct_accel(idx) + models.ext_accel(pot, pos[idx])
        e = np.linalg.norm(st.accelerations()[idx] - d, axis=1) / np.linalg.norm(d, axis=1)
        med[prec] = float(np.median(e))
        st.close()
    print("median total error vs direct sum", med, flush=True)
    assert med["mixed_group"] < 0.8 * med["fp64"]


def test_group_walk_is_deterministic_eager_and_graph(dwarf_case):
    ctx, pot, pos, vel, m = dwarf_case
    a = nb.NBodyState(ctx, pot, pos, vel, m, precision="mixed_group")
    b = nb.NBodyState(ctx, pot, pos, vel, m, precision="mixed_group")
    assert np.array_equal(a.accelerations(), b.accelerations())
    b.set_graph(True)
    for rep in range(3):
        assert a.steps(3) == 0
        assert b.steps(3) == 0
        pa, va, aa = a.bodies()
        pb, vb, ab = b.bodies()
        print("replay", rep, "identical", np.array_equal(pa, pb), flush=True)
        assert np.array_equal(pa, pb) and np.array_equal(va, vb) and np.array_equal(aa, ab)
    a.close()
    b.close()


@pytest.mark.parametrize("block,depth", [(64, 64), (128, 256), (256, 128)])
def test_group_walk_launch_shapes_agree(dwarf_case, block, depth):
    """Workgroup size only changes how waves share a CU; the list depth
    changes the FP32 partial sums' grouping (within FP32 rounding)."""
    ctx, pot, pos, vel, m = dwarf_case
    st = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed_group")
    base = st.accelerations()
    c0 = st.walk_counters()
    st.set_tune(block, depth)
    st.gravity()
    e = rel_err(st.accelerations(), base, 0)
    print(block, depth, "median", np.median(e), "max", e.max(), flush=True)
    assert st.walk_counters()[:3] == c0[:3]
    assert np.median(e) <= 1e-6 and e.max() <= 1e-4
    st.close()


def test_group_walk_zero_softening_and_massless():
    pos, vel, m = plummer_bodies(3000, seed=12)
    m = m.copy()
    m[::50] = 0.0                        # test particles: never in the tree
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=0.0, theta=0.5,
                             allowIncest=True)
    st = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed_group")
    a = st.accelerations()
    assert np.isfinite(a).all()
    ref = nb.NBodyState(ctx, None, pos, vel, m).accelerations()
    e = rel_err(a, ref)
    print("eps2=0 vs FP64 tree: median", np.median(e), "max", e.max(), flush=True)
    assert np.median(e) <= 1e-3
    st.close()


def test_group_walk_in_a_batch_matches_direct_sum():
    members = []
    for k, n in enumerate((5000, 64 * 37 + 5, 9000)):
        members.append(plummer_bodies(n, seed=60 + k, mass=8.0 + k, radius=0.2,
                                      position=(2.0 * k, -1.0, 0.5 * k)))
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=1e-4, theta=0.5)
    bt = nb.NBodyBatch(ctx, None, members, precision="mixed_group")
    assert bt.statuses() == [0, 0, 0]
    for k, (pos, vel, m) in enumerate(members):
        _, _, a = bt.sim_bodies(k)
        one = nb.NBodyState(ctx, None, pos, vel, m, precision="fp64")
        idx = np.arange(0, pos.shape[0], 5, dtype=np.int32)
        d = one.direct_accel(idx)
        eg = np.median(rel_err(a[idx], d, 0))
        ef = np.median(rel_err(one.accelerations()[idx], d, 0))
        print(f"sim {k}: group {eg:.3e} reference lists {ef:.3e}", flush=True)
        assert eg < ef
        one.close()
    bt.close()


@pytest.mark.parametrize("ratio", [0.05, 0.1])
def test_group_expansion_matches_its_restatement(dwarf_case, ratio):
    """Far cells expanded about each group's centre (nbhip_set_group_expansion):
    the near / far split is decided in FP64 by the same test as the numpy
    restatement, so the counts agree exactly; the accelerations (near cells
    per body in FP32, far cells as an FP32 second-order expansion) agree
    with the restatement's FP64 evaluation of the same expansion to FP32
    rounding."""
    ctx, pot, pos, vel, m = dwarf_case
    st = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed_group", prestep=False)
    st.set_group_expansion(ratio)
    st.gravity()
    gw = _model(st, ctx)
    groups = gw.groups()
    c = st.walk_counters5()
    want = gw.expanded_counts(ratio, groups)
    print(f"ratio {ratio}: gpu (bb, near cells, far pairs) {(c[0], c[1], c[4])} model {want}",
          flush=True)
    assert (c[0], c[1], c[4]) == want and c[4] > 0
    acc = st.walk_accel()
    rng = np.random.default_rng(3)
    errs = []
    for gi in rng.choice(len(groups), size=40, replace=False):
        a, b = groups[gi]
        errs.append(rel_err(acc[gw.order[a:b]], gw.expanded_forces(a, b, ratio), 0))
    e = np.concatenate(errs)
    print(f"FP32 vs FP64 expansion: median {np.median(e):.3e} p99 {np.percentile(e, 99):.3e}",
          flush=True)
    assert np.median(e) <= 1e-6 and np.percentile(e, 99) <= 1e-5
    # the expansion's own error against the per-body lists (same accepted cells)
    full = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed_group").walk_accel()
    d = rel_err(acc, full, 0)
    print(f"expanded vs per-body lists: median {np.median(d):.3e} p99 {np.percentile(d, 99):.3e}",
          flush=True)
    st.close()


def test_group_walk_results_do_not_depend_on_dispatch_order_or_sharing():
    """The first walk of a state orders its groups by nothing (no costs yet)
    and shares none out; later walks order them by the last walk's costs
    and share the costliest groups out among several waves, by whole
    sub-groups.  A dense clump with a sparse outskirt makes a few groups far
    costlier than the mean, so later walks do share; every walk's lists and
    accelerations must be the first walk's, bit for bit, and the lists the
    restatement's."""
    rng = np.random.default_rng(41)
    pos, vel, m = plummer_bodies(20000, seed=41, radius=0.05)
    far = rng.normal(size=(300, 3))
    far *= (rng.uniform(5.0, 40.0, 300) / np.linalg.norm(far, axis=1))[:, None]
    pos = np.concatenate([pos, far])
    vel = np.concatenate([vel, np.zeros((300, 3))])
    m = np.concatenate([m, np.full(300, m[0])])
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=1e-6, theta=0.5)
    st = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed_group", prestep=False)
    st.set_group_expansion(0.1)
    st.gravity()
    first = st.walk_accel()
    c_first = st.walk_counters5()
    for rep in range(4):
        st.gravity()
        a = st.walk_accel()
        c = st.walk_counters5()
        print("walk", rep + 2, "identical", np.array_equal(a, first), c[:2], c[4], flush=True)
        assert np.array_equal(a, first)
        assert c == c_first
    gw = _model(st, ctx)
    want = gw.expanded_counts(0.1, gw.groups())
    assert (c_first[0], c_first[1], c_first[4]) == want
    st.close()
