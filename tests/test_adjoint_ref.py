"""Pins the CPU adjoint reference (oracle/adjoint_ref.py) before it is used
to check the HIP adjoint: its forward map against the oracle's Exact
criterion run, the potential Jacobian against central differences of the
acceleration, and the full reverse sweep against central differences of the
forward map in positions, velocities, masses and the timestep."""
import numpy as np
import pytest

from helpers import plummer_bodies
from oracle import adjoint_ref as ar
from oracle import oracle as orc

POTS = {
    "mw": orc.milky_way(),
    "nfw_plummer_flat": orc.make_potential(sphere=(2, 3.0e4, 0.9), disk=(1, 2.0e5, 4.0, 0.5),
                                           disk2=(1, 5.0e4, 2.0, 0.1), halo=(2, 120.0, 15.0, 1.0)),
    "log_flattened": orc.make_potential(sphere=(1, 1.0e5, 0.5), halo=(1, 90.0, 10.0, 0.8)),
}


@pytest.mark.parametrize("name", sorted(POTS))
def test_numpy_ext_accel_matches_oracle(name):
    rng = np.random.default_rng(3)
    x = rng.normal(size=(40, 3)) * 20.0
    x[0] = (1e-3, -2e-3, 5e-4)                     # inside the r clamp
    p = ar.pot_params(POTS[name])
    got = ar.ext_accel(p, x)
    ref = orc.ext_accel(POTS[name], x)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("name", sorted(POTS))
def test_ext_jacobian_is_the_derivative(name):
    rng = np.random.default_rng(4)
    x = rng.normal(size=(30, 3)) * 15.0
    x[0] = (1e-3, -2e-3, 5e-4)
    p = ar.pot_params(POTS[name])
    g = rng.normal(size=x.shape)
    jg = ar.ext_jac(p, x, g)
    h = 1e-5
    fd = np.zeros_like(x)
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        col = (ar.ext_accel(p, x + e) - ar.ext_accel(p, x - e)) / (2 * h)   # column d of J
        fd[:, d] = (col * g).sum(1)                                         # (J^T g)_d
    err = np.abs(jg - fd).max(1) / np.abs(fd).max(1)
    assert err.max() < 1e-6, err.max()


def test_forward_matches_oracle_exact_run():
    pos, vel, m = plummer_bodies(64, seed=5)
    pot = orc.milky_way()
    pos = pos + (8.0, 3.0, -2.0)
    eps2, dt = 1e-3, 1e-3
    states = ar.kdk_forward(pos, vel, m, eps2, dt, 5, pot)
    o = orc.Oracle(orc.make_ctx(eps2=eps2, theta=0.0, timestep=dt, criterion="exact",
                                use_quad=False, potential=pot), pos, vel, m)
    assert o.run(5) == 0
    x, v, a = states[-1]
    np.testing.assert_allclose(x, o.pos, rtol=0, atol=1e-12 * np.abs(x).max())
    np.testing.assert_allclose(v, o.vel, rtol=0, atol=1e-11 * np.abs(v).max())
    np.testing.assert_allclose(a, o.acc, rtol=0, atol=1e-10 * np.abs(a).max())


@pytest.mark.parametrize("use_pot", [False, True])
def test_reverse_sweep_equals_finite_differences(use_pot):
    rng = np.random.default_rng(6)
    n, steps, eps2 = 12, 6, 4e-3
    pos, vel, m = plummer_bodies(n, seed=8)
    pos = pos + ((8.0, 3.0, -2.0) if use_pot else 0.0)
    pot = orc.milky_way() if use_pot else None
    dt = 2e-3
    wx, wv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))

    c0 = pos.mean(0)          # centred loss keeps the differences well above round-off

    def loss(p, v, mm, t):
        x, vv, _ = ar.kdk_forward(p, v, mm, eps2, t, steps, pot)[-1]
        return (wx * (x - c0)).sum() + (wv * vv).sum() + 0.5 * ((x - c0) ** 2).sum()

    states = ar.kdk_forward(pos, vel, m, eps2, dt, steps, pot)
    xN = states[-1][0]
    gx, gv, gm, gdt = ar.kdk_adjoint(states, m, eps2, dt, wx + (xN - c0), wv, pot)
    # central differences with steps small against the system's scales
    # (softening ~0.06, crossing time ~0.01) yet far above round-off
    def fd(fun, base, idx, step):
        b1, b2 = base.copy(), base.copy()
        b1[idx] += step
        b2[idx] -= step
        return (fun(b1) - fun(b2)) / (2 * step)

    for i, d in [(0, 0), (3, 1), (7, 2), (11, 0)]:
        want = fd(lambda p: loss(p, vel, m, dt), pos, (i, d), 1e-5)
        assert abs(gx[i, d] - want) <= 1e-6 * max(abs(want), 1e-3), ("x", i, d, gx[i, d], want)
        want = fd(lambda v: loss(pos, v, m, dt), vel, (i, d), 1e-4)
        assert abs(gv[i, d] - want) <= 1e-6 * max(abs(want), 1e-3), ("v", i, d, gv[i, d], want)
    for i in (0, 5, 9):
        want = fd(lambda mm: loss(pos, vel, mm, dt), m, i, 1e-4 * m[i])
        assert abs(gm[i] - want) <= 1e-6 * max(abs(want), 1e-3), ("m", i, gm[i], want)
    want = fd(lambda t: loss(pos, vel, m, t[0]), np.array([dt]), 0, 1e-4 * dt)
    assert abs(gdt - want) <= 1e-6 * abs(want), ("dt", gdt, want)


@pytest.mark.parametrize("criterion,quad", [("bh86", True), ("bh86", False), ("sw93", True)])
def test_frozen_tree_vjp_equals_differences(criterion, quad):
    """The frozen-tree force reproduces the oracle's tree walk, and its
    vector-Jacobian product (as the GPU computes it) equals central
    differences of that force along random directions in positions and
    masses; with the lists frozen the force is smooth, so they converge."""
    pos, vel, m = plummer_bodies(300, seed=21)
    m = m * np.random.default_rng(1).uniform(0.5, 1.5, m.shape)
    eps2 = 1e-4
    o = orc.Oracle(orc.make_ctx(eps2=eps2, theta=0.6, criterion=criterion, use_quad=quad),
                   pos, vel, m)
    assert o.gravity() == 0
    t = o.tree()
    ch = ar.tree_children(t)
    lists = ar.walk_lists(t)
    a = ar.frozen_accel(pos, m, t, ch, lists, eps2, quad)
    np.testing.assert_allclose(a, o.acc, rtol=0, atol=1e-12 * np.abs(o.acc).max())
    rng = np.random.default_rng(2)
    G = rng.normal(size=pos.shape)
    gx, gm = ar.frozen_vjp(pos, m, t, ch, lists, eps2, G, quad)
    # truncation error falls 100x per 10x step here (smooth frozen map): 1e-7 -> ~1e-9
    for _ in range(3):
        D = rng.normal(size=pos.shape)
        h = 1e-7
        fd = ((ar.frozen_accel(pos + h * D, m, t, ch, lists, eps2, quad)
               - ar.frozen_accel(pos - h * D, m, t, ch, lists, eps2, quad)) * G).sum() / (2 * h)
        assert abs((gx * D).sum() - fd) <= 1e-7 * abs(fd)
    Dm = rng.normal(size=m.shape) * m
    h = 1e-7
    fd = ((ar.frozen_accel(pos, m + h * Dm, t, ch, lists, eps2, quad)
           - ar.frozen_accel(pos, m - h * Dm, t, ch, lists, eps2, quad)) * G).sum() / (2 * h)
    assert abs((gm * Dm).sum() - fd) <= 1e-7 * abs(fd)


@pytest.mark.parametrize("use_pot", [False, True])
def test_frozen_tree_reverse_sweep_equals_differences(use_pot):
    """Several KDK steps of a theta = 0.6 tree run with each state's tree and
    lists held fixed: the frozen forward map reproduces the oracle's run, and
    its reverse sweep (frozen_vjp in each step) equals central differences of
    that map along random directions in x0, v0, m and for dt.  The map is
    smooth, so the differences converge (the gap falls ~100x per 10x step)."""
    n, steps, eps2, dt = 200, 5, 1e-4, 2e-4
    pos, vel, m = plummer_bodies(n, seed=31)
    m = m * np.random.default_rng(4).uniform(0.5, 1.5, n)
    pos = pos + ((8.0, 3.0, -2.0) if use_pot else 0.0)
    pot = orc.milky_way() if use_pot else None
    o = orc.Oracle(orc.make_ctx(eps2=eps2, theta=0.6, timestep=dt, potential=pot), pos, vel, m)
    assert o.gravity() == 0
    frozen = ar.frozen_trees(o, steps)
    states = ar.kdk_forward_frozen(pos, vel, m, eps2, dt, frozen, pot)
    x, v, a = states[-1]
    np.testing.assert_allclose(x, o.pos, rtol=0, atol=1e-12 * np.abs(x).max())
    np.testing.assert_allclose(v, o.vel, rtol=0, atol=1e-11 * np.abs(v).max())
    np.testing.assert_allclose(a, o.acc, rtol=0, atol=1e-10 * np.abs(a).max())

    rng = np.random.default_rng(5)
    c0 = pos.mean(0)
    wx, wv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))

    def loss(p, vv, mm, t):
        xN, vN, _ = ar.kdk_forward_frozen(p, vv, mm, eps2, t, frozen, pot)[-1]
        return (wx * (xN - c0)).sum() + (wv * vN).sum() + 0.5 * ((xN - c0) ** 2).sum()

    gx, gv, gm, gdt = ar.kdk_adjoint_frozen(states, m, eps2, dt, frozen, wx + (x - c0), wv, pot)
    D = rng.normal(size=(n, 3))
    h = 1e-7
    fd = (loss(pos + h * D, vel, m, dt) - loss(pos - h * D, vel, m, dt)) / (2 * h)
    assert abs((gx * D).sum() - fd) <= 1e-7 * abs(fd), ("x", (gx * D).sum(), fd)
    fd = (loss(pos, vel + h * D, m, dt) - loss(pos, vel - h * D, m, dt)) / (2 * h)
    assert abs((gv * D).sum() - fd) <= 1e-7 * abs(fd), ("v", (gv * D).sum(), fd)
    # masses: below h ~ 1e-6 the differences are round-off-limited here
    Dm = rng.normal(size=n) * m
    hm = 1e-5
    fd = (loss(pos, vel, m + hm * Dm, dt) - loss(pos, vel, m - hm * Dm, dt)) / (2 * hm)
    assert abs((gm * Dm).sum() - fd) <= 1e-7 * abs(fd), ("m", (gm * Dm).sum(), fd)
    hd = 1e-5 * dt
    fd = (loss(pos, vel, m, dt + hd) - loss(pos, vel, m, dt - hd)) / (2 * hd)
    assert abs(gdt - fd) <= 1e-7 * abs(fd), ("dt", gdt, fd)
