"""Many simulations in one state (nbhip_batch_*): every simulation of a batch
must evolve exactly as it would alone -- same tree, same accelerations, same
trajectory, bit for bit -- because its tree, walk order and sums do not
depend on the other simulations.  Also: graph-captured steps equal eager
steps, the walk autotuner's settings leave FP64 results unchanged, and a
failing simulation is reported on its own status word."""
import numpy as np
import pytest

from helpers import plummer_bodies
from milkywayathome_client_amd import nbody as nb
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 37, 1000, 5000, 20000]


def _members(sizes=SIZES):
    out = []
    for k, n in enumerate(sizes):
        pos, vel, m = plummer_bodies(n, seed=100 + k, mass=5.0 + k, radius=0.1 + 0.05 * k,
                                     position=(3.0 * k - 6.0, 1.5 * k, -2.0 * k),
                                     velocity=(10.0 * k, -5.0, 3.0))
        out.append((pos, vel, m))
    return out


def _eps2(sizes=SIZES):
    return [1e-4 * (1 + 0.5 * k) for k in range(len(sizes))]


@pytest.mark.parametrize("precision,criterion", [("fp64", "bh86"), ("fp64", "sw93"),
                                                 ("fp64", "exact"), ("mixed", "bh86")])
def test_batch_equals_individual_states(precision, criterion):
    members, eps2 = _members(), _eps2()
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=1e-4, theta=0.6,
                             criterion=criterion)
    pot = nb.milky_way_potential()
    print(f"batch of {len(members)} sims {SIZES}, {precision} {criterion}", flush=True)
    bt = nb.NBodyBatch(ctx, pot, members, eps2=eps2, precision=precision)
    bt.steps(4)
    for k, (pos, vel, m) in enumerate(members):
        ck = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=eps2[k], theta=0.6,
                                criterion=criterion)
        one = nb.NBodyState(ck, pot, pos, vel, m, precision=precision)
        one.steps(4)
        p1, v1, a1 = one.bodies()
        pb, vb, ab = bt.sim_bodies(k)
        assert np.array_equal(pb, p1) and np.array_equal(vb, v1) and np.array_equal(ab, a1), k
        if criterion != "exact":
            i1, ib = one.tree_info(), bt.sim_tree_info(k)
            assert (ib.rsize, ib.ncell, ib.nnode, ib.max_depth, ib.nmassive) == \
                (i1.rsize, i1.ncell, i1.nnode, i1.max_depth, i1.nmassive), k
        one.close()
    assert bt.statuses() == [0] * len(members)
    bt.close()


def test_batch_member_matches_oracle():
    members, eps2 = _members(), _eps2()
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=1e-4, theta=0.5)
    bt = nb.NBodyBatch(ctx, None, members, eps2=eps2)
    bt.steps(3)
    for k in (3, 5):
        pos, vel, m = members[k]
        o = orc.Oracle(orc.make_ctx(eps2=eps2[k], theta=0.5, timestep=ctx.timestep), pos, vel, m)
        assert o.gravity() == 0
        for _ in range(3):
            assert o.step() == 0
        p, v, a = bt.sim_bodies(k)
        assert np.array_equal(p, o.pos) and np.array_equal(v, o.vel) and np.array_equal(a, o.acc)
    bt.close()


@pytest.mark.parametrize("batch", [False, True])
def test_graph_steps_equal_eager_steps(batch):
    pot = nb.milky_way_potential()
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=1e-4, theta=0.5)
    if batch:
        members = _members([300, 4000, 9000])
        mk = lambda: nb.NBodyBatch(ctx, pot, members)
    else:
        pos, vel, m = plummer_bodies(30000, seed=7)
        mk = lambda: nb.NBodyState(ctx, pot, pos, vel, m)
    eager, graph = mk(), mk()
    graph.set_graph(True)
    for rep in range(3):           # replays with host work and syncs in between
        assert eager.steps(5) == 0
        assert graph.steps(5) == 0
        pe, ve, ae = eager.bodies()
        pg, vg, ag = graph.bodies()
        print("replay", rep, "identical", np.array_equal(pg, pe), flush=True)
        assert np.array_equal(pg, pe) and np.array_equal(vg, ve) and np.array_equal(ag, ae)
    assert graph.tree_info().step == eager.tree_info().step == 15
    eager.close()
    graph.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_steps_with_a_deep_tree(precision):
    """A close pair inside a Plummer sphere pushes the tree past depth 24,
    where the graph-captured step summarises the deep levels in one
    single-workgroup launch: graph steps still equal eager steps."""
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=1e-4, theta=0.5)
    pos, vel, m = plummer_bodies(3000, seed=9)
    pos = np.concatenate([pos, pos[:1] + np.array([[1e-7, -1e-7, 5e-8]])])
    vel = np.concatenate([vel, vel[:1]])
    m = np.concatenate([m, m[:1]])
    eager = nb.NBodyState(ctx, None, pos, vel, m, precision=precision)
    graph = nb.NBodyState(ctx, None, pos, vel, m, precision=precision)
    depth = eager.tree_info().max_depth
    print("max depth", depth, flush=True)
    assert depth >= 26
    graph.set_graph(True)
    for rep in range(3):
        assert eager.steps(4) == 0
        assert graph.steps(4) == 0
        pe, ve, ae = eager.bodies()
        pg, vg, ag = graph.bodies()
        assert np.array_equal(pg, pe) and np.array_equal(vg, ve) and np.array_equal(ag, ae)
    assert graph.tree_info().max_depth >= 26
    eager.close()
    graph.close()


def test_autotune_settings_keep_fp64_results():
    pos, vel, m = plummer_bodies(50000, seed=11)
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=1e-4, theta=0.5)
    st = nb.NBodyState(ctx, None, pos, vel, m)
    a0 = st.walk_accel()
    best, trials = st.autotune()
    print("best", best, "trials", trials)
    assert len(trials) >= 3 and best[0] in (64, 128, 256) and best[2] > 0.0
    for block in (64, 128, 256):
        st.set_tune(block, 1024)
        assert np.array_equal(st.walk_accel(), a0), block
    st.close()


def test_autotune_mixed_within_tolerance():
    pos, vel, m = plummer_bodies(50000, seed=12)
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=1e-4, theta=0.5)
    ref = nb.NBodyState(ctx, None, pos, vel, m).walk_accel()
    st = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed")
    best, trials = st.autotune()
    print("best", best, "trials", len(trials))
    assert len(trials) >= 12
    for block, depth in ((64, 256), (256, 16384), (best[0], best[1])):
        st.set_tune(block, depth)
        a = st.walk_accel()
        e = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
        assert np.median(e) <= 1e-6, (block, depth, np.median(e))
    # partial sums follow each body's own node blocks: the workgroup size
    # does not change the mixed result either
    st.set_tune(64, 1024)
    a64 = st.walk_accel()
    st.set_tune(256, 1024)
    assert np.array_equal(st.walk_accel(), a64)
    st.close()


def test_coincident_bodies_flag_only_their_simulation():
    members = _members([500, 400, 600])
    pos, vel, m = members[1]
    pos = pos.copy()
    pos[7] = pos[3]                       # two bodies at one point
    members[1] = (pos, vel, m)
    octx = orc.make_ctx(eps2=1e-4, theta=0.5)
    assert orc.Oracle(octx, pos, vel, m).gravity() == nb.NBODY_TREE_STRUCTURE_ERROR
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=1e-4, theta=0.5)
    bt = nb.NBodyBatch(ctx, None, members, prestep=False)
    rc = bt.lib.nbhip_gravity(bt._h)
    st = bt.statuses()
    print("rc", hex(rc), "per sim", [hex(x) for x in st])
    assert rc & nb.NBODY_TREE_STRUCTURE_ERROR
    assert st[0] == 0 and st[2] == 0 and st[1] & nb.NBODY_TREE_STRUCTURE_ERROR
    assert nb.status_is_fatal(rc)
    assert bt.lib.nbhip_steps(bt._h, 1) == rc      # no further steps once fatal
    bt.close()
    with pytest.raises(nb.NBodyError) as e:
        nb.NBodyState(ctx, None, pos, vel, m)
    assert e.value.status & nb.NBODY_TREE_STRUCTURE_ERROR


def test_coherent_batch_refines_each_simulation():
    """The wave-coherent walk in a batch, where waves hold bodies of several
    simulations (the 1-, 2- and 37-body members share one wave with the start
    of the next): each simulation's error against its own FP64 direct sum is
    no larger than its FP64 tree's, the one- and two-body members get their
    exact forces, results do not depend on the workgroup size, and every
    status word stays clean."""
    members, eps2 = _members(), _eps2()
    ctx = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=1e-4, theta=0.6)
    print(f"coherent batch of {len(members)} sims {SIZES}", flush=True)
    bt = nb.NBodyBatch(ctx, None, members, eps2=eps2, precision="mixed_coherent")
    assert bt.statuses() == [0] * len(members)
    for k, (pos, vel, m) in enumerate(members):
        ck = nb.NBodyCtx.create(timestep=1e-4, timeEvolve=1e-2, eps2=eps2[k], theta=0.6)
        one = nb.NBodyState(ck, None, pos, vel, m)
        ab = bt.sim_bodies(k)[2]
        a64 = one.accelerations()
        if len(m) <= 2:
            np.testing.assert_allclose(ab, a64, rtol=1e-6, atol=1e-12)
        else:
            idx = np.arange(0, len(m), max(1, len(m) // 2000), dtype=np.int32)
            d = one.direct_accel(idx)
            e64 = np.median(np.linalg.norm(a64[idx] - d, axis=1) / np.linalg.norm(d, axis=1))
            ecoh = np.median(np.linalg.norm(ab[idx] - d, axis=1) / np.linalg.norm(d, axis=1))
            print(k, len(m), "median vs direct: fp64 tree", e64, "coherent batch", ecoh)
            assert ecoh <= 1.02 * e64, k
        one.close()
    bt.set_tune(64, 1024)
    a = bt.walk_accel()
    bt.set_tune(256, 1024)
    assert np.array_equal(bt.walk_accel(), a)
    bt.close()
