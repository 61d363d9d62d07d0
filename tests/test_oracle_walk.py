"""Pins the oracle's walk (nbGravity, nbody/src/nbody_grav.c:42-124) and
direct sum (nbGravity_Exact, :205-227) against closed forms and identities."""
import numpy as np
import pytest

from helpers import plummer_bodies, rel_err
from oracle import oracle as orc


def _oracle(pos, mass, **kw):
    ctx = orc.make_ctx(**kw)
    o = orc.Oracle(ctx, pos, np.zeros_like(pos), mass)
    rc = o.gravity()
    return o, rc


def test_two_body_closed_form():
    pos = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, 1.2]])
    mass = np.array([2.0, 5.0])
    eps2 = 1e-3
    o, rc = _oracle(pos, mass, eps2=eps2, theta=0.5)
    assert rc == 0
    d = pos[1] - pos[0]
    r2 = d @ d + eps2
    np.testing.assert_allclose(o.acc[0], mass[1] * d / r2 ** 1.5, rtol=1e-15)
    np.testing.assert_allclose(o.acc[1], -mass[0] * d / r2 ** 1.5, rtol=1e-15)


def test_direct_sum_conserves_momentum():
    pos, _, m = plummer_bodies(500)
    a = orc.direct_accel(pos, m, 1e-4)
    f = (m[:, None] * a).sum(0)
    assert np.abs(f).max() < 1e-12 * np.abs(m[:, None] * a).sum()


@pytest.mark.parametrize("criterion", ["bh86", "sw93", "treecode"])
def test_theta_zero_opens_everything(criterion):
    """findRCrit with theta = 0 sets rcrit2 = (2 rsize)^2 (nbody_tree.c:312-317):
    every cell is opened and the walk is the direct sum up to summation order."""
    pos, _, m = plummer_bodies(1500)
    o, rc = _oracle(pos, m, eps2=1e-4, theta=0.0, criterion=criterion)
    assert rc == 0
    d = orc.direct_accel(pos, m, 1e-4)
    assert rel_err(o.acc, d, 0).max() < 5e-13


def test_quadrupole_far_field_of_a_point_pair():
    """A cell holding two point masses seen from far away: the monopole +
    quadrupole term reproduces the exact pair force to O((s/r)^3); without
    the quadrupole the error is O((s/r)^2)."""
    s = 0.01
    pos = np.array([[0.3 + s, 0.21, 0.1], [0.3 - s, 0.19, 0.13], [1.9, 1.7, 1.8]])
    mass = np.array([1.0, 2.0, 1e-9])
    d = pos[:2] - pos[2]
    exact = (mass[:2, None] * d / np.linalg.norm(d, axis=1)[:, None] ** 3).sum(0)
    errs = {}
    for quad in (True, False):
        # theta = 1.5: body 2 accepts the root cell, whose moments are those of
        # the pair (body 2 itself weighs 1e-9)
        o, _ = _oracle(pos, mass, eps2=0.0, theta=1.5, use_quad=quad, allow_incest=True)
        errs[quad] = np.linalg.norm(o.acc[2] - exact) / np.linalg.norm(exact)
    r = np.linalg.norm(pos[2] - pos[:2].mean(0))
    assert errs[True] < 0.3 * (s / r) ** 2
    assert errs[False] > 10 * errs[True]


def test_quadrupole_improves_plummer_accuracy():
    pos, _, m = plummer_bodies(3000)
    d = orc.direct_accel(pos, m, 1e-4)
    e = {}
    for quad in (True, False):
        o, _ = _oracle(pos, m, eps2=1e-4, theta=0.7, use_quad=quad)
        e[quad] = np.median(rel_err(o.acc, d, 0))
    assert e[True] < 0.5 * e[False]


def test_error_falls_with_theta():
    pos, _, m = plummer_bodies(3000, seed=3)
    d = orc.direct_accel(pos, m, 1e-4)
    errs = []
    for th in (1.0, 0.7, 0.5, 0.3):
        o, _ = _oracle(pos, m, eps2=1e-4, theta=th, allow_incest=True)
        errs.append(np.median(rel_err(o.acc, d, 0)))
    assert all(b < a for a, b in zip(errs, errs[1:]))


def test_incest_with_wide_opening_angle():
    """BH86 with theta >= 1/sqrt(3)... can accept a cell containing the body
    itself; the body then never meets itself (nbody_grav.c:116-121)."""
    pos, _, m = plummer_bodies(3000)
    o, rc = _oracle(pos, m, eps2=1e-4, theta=1.2, allow_incest=True)
    assert rc == 1 << 24                          # NBODY_TREE_INCEST_NONFATAL
    o, rc = _oracle(pos, m, eps2=1e-4, theta=1.2, allow_incest=False)
    assert rc == 1 << 4                           # NBODY_TREE_INCEST_FATAL
    o, rc = _oracle(pos, m, eps2=1e-4, theta=0.5)
    assert rc == 0


def test_walk_counts_are_consistent():
    pos, _, m = plummer_bodies(2000)
    o, _ = _oracle(pos, m, eps2=1e-4, theta=0.5)
    nbb, nbc, nop = o.walk_counts(np.arange(2000))
    ncell, _, nnode = o.tree_summary()
    assert nbb > 0 and nbc > 0 and nop > 0
    # theta = 0: every body meets every other body, nothing else
    o0, _ = _oracle(pos, m, eps2=1e-4, theta=0.0)
    nbb0, nbc0, nop0 = o0.walk_counts(np.arange(2000))
    assert nbb0 == 2000 * 1999 and nbc0 == 0 and nop0 == 2000 * ncell
