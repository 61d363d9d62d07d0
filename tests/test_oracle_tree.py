"""Pins the oracle's tree (restating nbody/src/nbody_tree.c) against an
independent top-down construction from the definition: the root is centred
on the origin with side rsize (doubled until rsize >= 2 max|x|), a cell
exists wherever two or more massive bodies share a sub-cube, the octant bit
is set when midpoint <= coordinate, and per-cell sums are taken directly
over the bodies of the cell (not recursively)."""
import math

import numpy as np
import pytest

from oracle import oracle as orc


def brute_tree(pos, mass, rsize0):
    """Top-down partition; returns rsize and the pre-order node list."""
    xyzmax = max(np.abs(pos).max(), 0.0)
    rsize = rsize0
    while rsize < 2.0 * xyzmax:
        rsize *= 2.0
    nodes = []

    def visit(idx, mid, side, level):
        me = {"kind": 0, "level": level, "mid": mid, "bodies": idx, "side": side}
        nodes.append(me)
        groups = {}
        for b in idx:
            o = 4 * (mid[0] <= pos[b, 0]) + 2 * (mid[1] <= pos[b, 1]) + (mid[2] <= pos[b, 2])
            groups.setdefault(o, []).append(b)
        for o in sorted(groups):
            g = groups[o]
            if len(g) == 1:
                nodes.append({"kind": 1, "level": -1, "body": g[0]})
                continue
            q = 0.25 * side
            cm = (mid[0] + (q if o & 4 else -q), mid[1] + (q if o & 2 else -q),
                  mid[2] + (q if o & 1 else -q))
            visit(g, cm, 0.5 * side, level + 1)
        me["end"] = len(nodes)

    massive = [i for i in range(len(mass)) if mass[i] != 0.0]
    visit(massive, (0.0, 0.0, 0.0), rsize, 0)
    return rsize, nodes


def oracle_tree(pos, mass, theta=0.5, criterion="bh86", rsize0=4.0):
    ctx = orc.make_ctx(eps2=1e-4, theta=theta, criterion=criterion, tree_rsize=rsize0)
    o = orc.Oracle(ctx, pos, np.zeros_like(pos), mass)
    rc = o.gravity()
    return o, rc


def cluster(n, seed, scale=1.0, offset=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * scale * rng.uniform(0.05, 1.0, size=(n, 1)) + offset
    mass = rng.uniform(0.5, 2.0, size=n)
    return pos, mass


@pytest.mark.parametrize("n,seed,scale,offset", [(2, 0, 1.0, (0, 0, 0)), (40, 1, 0.3, (0, 0, 0)),
                                                 (300, 2, 2.0, (1.5, -7.0, 3.0)),
                                                 (500, 3, 0.05, (20.0, 30.0, -10.0))])
def test_topology_matches_definition(n, seed, scale, offset):
    pos, mass = cluster(n, seed, scale, offset)
    rsize, nodes = brute_tree(pos, mass, 4.0)
    o, rc = oracle_tree(pos, mass)
    assert rc == 0
    assert o.rsize == rsize
    t = o.tree()
    assert len(t["kind"]) == len(nodes)
    ncell = sum(1 for x in nodes if x["kind"] == 0)
    assert o.tree_summary()[0] == ncell
    for i, x in enumerate(nodes):
        assert t["kind"][i] == x["kind"]
        if x["kind"] == 0:
            assert t["level"][i] == x["level"]
            assert tuple(t["mid"][i]) == x["mid"]          # exact dyadic midpoints
            assert t["skip"][i] == x["end"]
        else:
            assert t["ref"][i] == x["body"]
            assert t["skip"][i] == i + 1


def test_max_depth_is_deepest_body_level():
    pos, mass = cluster(200, 7, 0.5)
    _, nodes = brute_tree(pos, mass, 4.0)
    o, _ = oracle_tree(pos, mass)
    deepest = max(x["level"] for x in nodes if x["kind"] == 0)
    assert o.tree_summary()[1] == deepest


@pytest.mark.parametrize("criterion,theta", [("bh86", 0.5), ("sw93", 0.7), ("treecode", 0.6),
                                             ("bh86", 0.0)])
def test_cell_moments_and_critical_radii(criterion, theta):
    pos, mass = cluster(250, 11, 0.8, (3.0, -2.0, 1.0))
    rsize, nodes = brute_tree(pos, mass, 4.0)
    o, rc = oracle_tree(pos, mass, theta=theta, criterion=criterion)
    t = o.tree()
    for i, x in enumerate(nodes):
        if x["kind"] != 0:
            continue
        b = np.array(x["bodies"])
        m = mass[b]
        M = math.fsum(m)
        cm = np.array([math.fsum(m * pos[b, d]) / M for d in range(3)])
        np.testing.assert_allclose(t["mass"][i], M, rtol=1e-15)
        np.testing.assert_allclose(t["pos"][i], cm, rtol=0, atol=1e-14 * (1 + np.abs(cm).max()))
        d = pos[b] - t["pos"][i]
        r2 = (d * d).sum(1)
        Q = np.array([np.sum(m * (3 * d[:, 0] * d[:, 0] - r2)), np.sum(m * 3 * d[:, 0] * d[:, 1]),
                      np.sum(m * 3 * d[:, 0] * d[:, 2]), np.sum(m * (3 * d[:, 1] * d[:, 1] - r2)),
                      np.sum(m * 3 * d[:, 1] * d[:, 2]), np.sum(m * (3 * d[:, 2] * d[:, 2] - r2))])
        scale = np.sum(m * r2) + 1e-300
        assert np.abs(t["quad"][i] - Q).max() <= 1e-12 * scale
        side = x["side"]
        mid = np.array(x["mid"])
        cmo = t["pos"][i]
        if theta == 0.0:
            want = (2.0 * rsize) ** 2
        elif criterion == "bh86":
            want = (side / theta) ** 2
        elif criterion == "treecode":
            want = (side / theta + np.linalg.norm(cmo - mid)) ** 2
        else:
            dmin = cmo - (mid - 0.5 * side)
            want = np.sum(np.maximum(dmin, side - dmin) ** 2) / theta ** 2
        np.testing.assert_allclose(t["rcrit2"][i], want, rtol=1e-14)


def test_massless_bodies_not_inserted():
    pos, mass = cluster(100, 5, 0.4)
    mass[::3] = 0.0
    rsize, nodes = brute_tree(pos, mass, 4.0)
    o, _ = oracle_tree(pos, mass)
    t = o.tree()
    assert len(t["kind"]) == len(nodes)
    assert set(t["ref"][t["kind"] == 1]) == set(np.nonzero(mass)[0])


def test_root_doubles_and_persists():
    pos, mass = cluster(50, 9, 0.2)
    pos[0] = (40.0, 0.0, 0.0)
    o, _ = oracle_tree(pos, mass)
    assert o.rsize == 128.0
    # pull the outlier back in: rsize never decreases (nbody_tree.c:173-175)
    for d in range(3):
        o.st.contents.pos[d] = 0.0
    o.gravity()
    assert o.rsize == 128.0


def test_chain_cells_for_close_pair():
    pos = np.array([[0.1, 0.1, 0.1], [0.1 + 1e-9, 0.1, 0.1], [-1.0, 0.5, 0.2]])
    mass = np.ones(3)
    rsize, nodes = brute_tree(pos, mass, 4.0)
    o, rc = oracle_tree(pos, mass)
    assert rc == 0
    deepest = max(x["level"] for x in nodes if x["kind"] == 0)
    assert deepest > 20                   # one chain cell per level down to the split
    assert o.tree_summary()[0] == sum(1 for x in nodes if x["kind"] == 0)
    assert o.tree_summary()[1] == deepest


def test_coincident_bodies_are_a_structure_error():
    pos = np.array([[0.3, 0.2, 0.1], [0.3, 0.2, 0.1], [1.0, 1.0, 1.0]])
    _, rc = oracle_tree(pos, np.ones(3))
    assert rc == 8   # NBODY_TREE_STRUCTURE_ERROR
