"""Trees deeper than the sort word, against the oracle node for node.

The radix sort orders bodies by the top 21 Morton levels (19 for a batch of
64 simulations: the simulation id takes 6 key bits).  Bodies that agree on
those levels are ordered by levels 22-63 afterwards (k_tiefix), the Karras
build reads the 189-bit prefixes past the sort word, and a graph-captured
step summarises levels 24-63 in one single-workgroup launch.  Every case
here builds cells below the sort word and compares the HIP tree (kind,
Next, level, centre of mass, mass, critical radius, quadrupole), the
accelerations and a few KDK steps with the oracle's serial insertion
(nbody/src/nbody_tree.c:36-52,242-282), in eager and graph-captured steps.
The FP64 form must be bit-identical; the mixed form keeps the FP64 tree
(compared bit for bit) and walks it in FP32.
"""
import numpy as np
import pytest

from helpers import plummer_bodies, rel_err
from milkywayathome_client_amd import nbody as nb
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

EPS2 = 1e-4
DT = 1e-4


def _close_pair(n=3000, seed=9, offset=(1e-7, -1e-7, 5e-8)):
    """A Plummer sphere plus a copy of its first body displaced by ~1e-7:
    the pair shares about 25 levels (root side 4)."""
    pos, vel, m = plummer_bodies(n, seed=seed)
    pos = np.concatenate([pos, pos[:1] + np.array([offset])])
    vel = np.concatenate([vel, vel[:1]])
    m = np.concatenate([m, m[:1]])
    return pos, vel, m


def _cluster(n=3000, nclust=300, seed=5):
    """A Plummer sphere plus `nclust` bodies spread uniformly over a cube of
    side 4 * 2^-23 around one point: they all share the first 22 levels, so
    one run of equal sort words holds all of them."""
    pos, vel, m = plummer_bodies(n, seed=seed)
    rng = np.random.default_rng(seed)
    side = 4.0 * 2.0 ** -23
    corner = np.array([0.0625 + 3 * side, -0.03125 + 5 * side, 0.015625 + 7 * side])
    cl = corner + rng.uniform(0.0, side, size=(nclust, 3))
    pos = np.concatenate([pos, cl])
    vel = np.concatenate([vel, np.zeros((nclust, 3))])
    m = np.concatenate([m, np.full(nclust, m[0])])
    return pos, vel, m


def _boundary(n=2000, seed=3):
    """Bodies exactly on the faces and corners of the root (|x| = rsize/2 =
    2 with treeRSize 4, the largest the root-size rule keeps without
    doubling): the reference's subIndex puts x = +2 on the upper side of
    every midpoint and x = -2 on the lower side."""
    pos, vel, m = plummer_bodies(n, seed=seed)
    pos = np.clip(pos, -1.9, 1.9)
    extra = np.array([[2.0, 0.3, -0.5], [2.0, 0.31, -0.5], [-2.0, 1.0, 2.0], [2.0, 2.0, 2.0],
                      [-2.0, -2.0, -2.0], [1.5, -2.0, 2.0], [2.0, 2.0, 1.999999999999]])
    pos = np.concatenate([pos, extra])
    vel = np.concatenate([vel, np.zeros_like(extra)])
    m = np.concatenate([m, np.full(extra.shape[0], m[0])])
    return pos, vel, m


CASES = {"close_pair": _close_pair, "cluster": _cluster, "boundary": _boundary}


def _oracle(bodies, eps2=EPS2, theta=0.5):
    pos, vel, m = bodies
    return orc.Oracle(orc.make_ctx(eps2=eps2, theta=theta, timestep=DT), pos, vel, m)


def _assert_same_tree(gt, ct, label):
    assert gt["kind"].shape == ct["kind"].shape, (label, gt["kind"].shape, ct["kind"].shape)
    for k in ("kind", "skip", "level"):
        assert np.array_equal(gt[k], ct[k]), (label, k)
    bodies = ct["kind"] == 1
    assert np.array_equal(gt["ref"][bodies], ct["ref"][bodies]), (label, "ref")
    for k in ("pos", "mass", "rcrit2", "quad"):
        np.testing.assert_array_equal(gt[k], ct[k], err_msg=f"{label} {k}")


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("graph", [False, True])
def test_deep_tree_fp64_matches_oracle(case, graph):
    bodies = CASES[case]()
    pos, vel, m = bodies
    ctx = nb.NBodyCtx.create(timestep=DT, timeEvolve=1e-2, eps2=EPS2, theta=0.5)
    print(f"case {case} graph {graph} n {pos.shape[0]}", flush=True)
    gpu = nb.NBodyState(ctx, None, pos, vel, m)
    cpu = _oracle(bodies)
    assert cpu.gravity() == 0
    ncell, depth, nnode = cpu.tree_summary()
    info = gpu.tree_info()
    print("depth", depth, "ncell", ncell, "gpu depth", info.max_depth, flush=True)
    if case != "boundary":
        assert depth >= 24
    assert (info.ncell, info.nnode, info.max_depth, info.rsize) == (ncell, nnode, depth, cpu.rsize)
    _assert_same_tree(gpu.tree(), cpu.tree(), f"{case} pre-step")
    assert np.array_equal(gpu.accelerations(), cpu.acc)
    gpu.set_graph(graph)
    for rep in range(2):
        assert gpu.steps(3) == 0
        for _ in range(3):
            assert cpu.step() == 0
        p, v, a = gpu.bodies()
        assert np.array_equal(p, cpu.pos) and np.array_equal(v, cpu.vel), (case, rep)
        assert np.array_equal(a, cpu.acc), (case, rep)
        _assert_same_tree(gpu.tree(), cpu.tree(), f"{case} after {3 * (rep + 1)} steps")
    gpu.close()


@pytest.mark.parametrize("case", ["close_pair", "cluster"])
@pytest.mark.parametrize("graph", [False, True])
def test_deep_tree_mixed_keeps_the_fp64_tree(case, graph):
    bodies = CASES[case]()
    pos, vel, m = bodies
    ctx = nb.NBodyCtx.create(timestep=DT, timeEvolve=1e-2, eps2=EPS2, theta=0.5)
    print(f"case {case} mixed graph {graph}", flush=True)
    gpu = nb.NBodyState(ctx, None, pos, vel, m, precision="mixed")
    cpu = _oracle(bodies)
    assert cpu.gravity() == 0
    _assert_same_tree(gpu.tree(), cpu.tree(), f"{case} mixed")
    e = rel_err(gpu.accelerations(), cpu.acc, 0)
    print("mixed vs oracle: median", np.median(e), "max", e.max(), flush=True)
    assert np.median(e) <= 1e-6 and np.percentile(e, 99) <= 1e-5
    gpu.set_graph(graph)
    assert gpu.steps(4) == 0
    for _ in range(4):
        assert cpu.step() == 0
    # the tree is built from the mixed run's own positions: same cells as the
    # oracle's wherever the positions agree to the cell boundaries
    info = gpu.tree_info()
    assert abs(info.max_depth - cpu.tree_summary()[1]) <= 1
    dx = np.abs(gpu.bodies()[0] - cpu.pos).max()
    print("mixed 4 steps max |dx|", dx, flush=True)
    assert dx < 1e-9
    gpu.close()


def _batch_members(nsim=64, n=150):
    out = []
    for k in range(nsim):
        pos, vel, m = plummer_bodies(n, seed=200 + k, mass=3.0 + 0.1 * k,
                                     radius=0.05 + 0.002 * k,
                                     position=(0.01 * k - 0.3, 0.2 - 0.005 * k, 0.003 * k),
                                     velocity=(1.0, -2.0, 0.5 * k))
        # a close pair in every simulation: ties past the 19-level sort word
        off = np.array([[1e-7 * (1 + k % 5), -3e-8, 2e-8 * (k % 3)]])
        pos = np.concatenate([pos, pos[k % n: k % n + 1] + off])
        vel = np.concatenate([vel, vel[k % n: k % n + 1]])
        m = np.concatenate([m, m[:1]])
        out.append((pos, vel, m))
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_deep_trees_in_a_64_simulation_batch_match_oracle(graph):
    members = _batch_members()
    eps2 = [EPS2 * (1 + 0.01 * k) for k in range(len(members))]
    ctx = nb.NBodyCtx.create(timestep=DT, timeEvolve=1e-2, eps2=EPS2, theta=0.5)
    print(f"batch of {len(members)} deep trees, graph {graph}", flush=True)
    bt = nb.NBodyBatch(ctx, None, members, eps2=eps2)
    oracles = [_oracle(b, eps2=e) for b, e in zip(members, eps2)]
    for o in oracles:
        assert o.gravity() == 0
    deepest = 0
    for k, o in enumerate(oracles):
        ncell, depth, nnode = o.tree_summary()
        ti = bt.sim_tree_info(k)
        assert (ti.ncell, ti.nnode, ti.max_depth, ti.rsize) == (ncell, nnode, depth, o.rsize), k
        _assert_same_tree(bt.tree(k), o.tree(), f"sim {k}")
        deepest = max(deepest, depth)
    print("deepest tree", deepest, flush=True)
    assert deepest >= 22
    bt.set_graph(graph)
    assert bt.steps(3) == 0
    for k, o in enumerate(oracles):
        for _ in range(3):
            assert o.step() == 0
        p, v, a = bt.sim_bodies(k)
        assert np.array_equal(p, o.pos) and np.array_equal(v, o.vel), k
        assert np.array_equal(a, o.acc), k
    for k in (0, 31, 63):
        _assert_same_tree(bt.tree(k), oracles[k].tree(), f"sim {k} after 3 steps")
    assert bt.statuses() == [0] * len(members)
    bt.close()
