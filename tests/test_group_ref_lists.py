"""The group walk's lists, restated in oracle/group_ref.py (the checker of the
GPU's group walk), against the reference's own opening test on the oracle's
tree: every node a sub-group accepts passes each member body's test
(|d|^2 >= rcrit2, nbody/src/nbody_grav.c:42-124), and the accepted nodes of
a sub-group tile the tree (each body of the tree is reached through exactly
one accepted node), so each member's list refines its reference list.  Also
the sub-group cuts: runs with a link at a sixteenth of the root or shallower,
or one more than two levels shallower than the run's median, are cut there."""
import numpy as np
import pytest

from helpers import plummer_bodies
from oracle import group_ref as gr
from oracle import oracle as orc


def _tree(n, seed, scale):
    pos, vel, m = plummer_bodies(n, seed=seed, radius=0.2)
    rng = np.random.default_rng(seed)
    far = rng.normal(size=(40, 3))
    far *= (rng.uniform(2.0, 12.0, 40) / np.linalg.norm(far, axis=1))[:, None]
    pos = np.concatenate([pos, far])
    vel = np.concatenate([vel, np.zeros((40, 3))])
    m = np.concatenate([m, np.full(40, m[0])])
    o = orc.Oracle(orc.make_ctx(eps2=1e-5, theta=0.5), pos, vel, m)
    o.gravity()
    tree = o.tree()
    return tree, gr.GroupWalk(tree, o.rsize, 1e-5, open_scale=scale)


def _leaves_under(kids, kind, node):
    out, stack = [], [node]
    while stack:
        q = stack.pop()
        if kind[q] == 1:
            out.append(q)
        else:
            stack.extend(int(c) for c in kids[q] if c >= 0)
    return out


@pytest.mark.parametrize("scale", [1.0, 1.45])
def test_accepted_nodes_pass_every_members_test_and_tile_the_tree(scale):
    tree, gw = _tree(3000, 17, scale)
    groups = gw.groups()
    rng = np.random.default_rng(5)
    nleaf = int((tree["kind"] == 1).sum())
    for gi in rng.choice(len(groups), size=25, replace=False):
        a, b = groups[gi]
        lst = gw.accepted(a, b)
        cells = lst[tree["kind"][lst] == 0]
        for s in range(a, b):
            d = tree["pos"][cells] - gw.spos[s]
            d2 = d[:, 2] * d[:, 2] + (d[:, 1] * d[:, 1] + d[:, 0] * d[:, 0])
            assert (d2 >= tree["rcrit2"][cells]).all()
        reached = np.concatenate([_leaves_under(gw.kids, tree["kind"], int(q)) for q in lst])
        assert reached.size == nleaf and np.unique(reached).size == nleaf


def test_cuts_follow_the_rules():
    # a run of 64 links at level 12 with one link at level 9: more than two
    # levels below the median, so the wave is cut after that lane
    lv = np.full(64, 12)
    lv[63] = -1
    lv[20] = 9
    assert gr.wave_groups(lv, 64) == [(0, 21), (21, 64)]
    # one level below the median: no cut
    lv[20] = 10
    assert gr.wave_groups(lv, 64) == [(0, 64)]
    # a link at level gr.SPLIT_ABS or shallower is cut whatever the median
    # (equal links: each pass cuts each run after its first lane)
    lv = np.full(64, 3)
    lv[63] = -1
    assert gr.wave_groups(lv, 64) == [(0, 1), (1, 2), (2, 3), (3, 64)]
    lv = np.full(64, gr.SPLIT_ABS + 1)
    lv[63] = -1
    assert gr.wave_groups(lv, 64) == [(0, 64)]
    # hard cuts (-1) always split
    lv = np.full(64, 12)
    lv[31] = -1
    lv[63] = -1
    assert gr.wave_groups(lv, 64) == [(0, 32), (32, 64)]
