"""HIP adjoint leapfrog (nbhip_adjoint_forward / nbhip_adjoint_backward).

Pinned against: the CPU reverse sweep of oracle/adjoint_ref.py (itself
checked against finite differences) for the Exact criterion; the Exact
adjoint again for the tree walk at theta = 0; the numpy frozen-tree
vector-Jacobian product on the oracle's identical tree for one force
evaluation at theta = 0.6 (central differences of the real tree run cannot
serve: moving bodies flips some of the ~10^7 opening tests and adds jumps
that the frozen-tree derivative by definition leaves out); individual runs
for batched simulations; and the forward run itself for the checkpointed
forward pass."""
import numpy as np
import pytest

from helpers import plummer_bodies
from milkywayathome_client_amd import nbody as nb
from oracle import adjoint_ref as ar
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _dwarf(n, seed, offset=(8.0, 3.0, -2.0), vel=(-40.0, 60.0, 30.0)):
    return plummer_bodies(n, seed=seed, mass=12.0, radius=0.2, position=offset, velocity=vel)


def _ctx(dt, eps2, criterion="bh86", theta=0.5, nsteps=10):
    return nb.NBodyCtx.create(timestep=dt, timeEvolve=dt * nsteps, eps2=eps2, theta=theta,
                              criterion=criterion)


def _loss_grads(x, v, c0, wx, wv):
    """L = wx.(x - c0) + wv.v + |x - c0|^2 / 2 on the final state."""
    return wx + (x - c0), wv


def test_checkpointed_forward_equals_run():
    pos, vel, m = _dwarf(3000, 1)
    ctx = _ctx(1e-3, 1e-3)
    pot = nb.milky_way_potential()
    a = nb.NBodyState(ctx, pot, pos, vel, m, prestep=False)
    a.adjoint_forward(13, 4)
    b = nb.NBodyState(ctx, pot, pos, vel, m, prestep=False)
    b.run(13)
    for u, w in zip(a.bodies(), b.bodies()):
        assert np.array_equal(u, w)


@pytest.mark.parametrize("use_pot", [False, True])
def test_exact_adjoint_matches_cpu_reverse_sweep(use_pot):
    n, steps, every, eps2, dt = 150, 17, 5, 4e-3, 1e-3
    pos, vel, m = _dwarf(n, 2) if use_pot else plummer_bodies(n, seed=2)
    rng = np.random.default_rng(9)
    wx, wv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    c0 = pos.mean(0)
    pot = nb.milky_way_potential() if use_pot else None
    st = nb.NBodyState(_ctx(dt, eps2, "exact"), pot, pos, vel, m, prestep=False)
    st.adjoint_forward(steps, every)
    x, v, _ = st.bodies()
    gxN, gvN = _loss_grads(x, v, c0, wx, wv)
    gx, gv, gm, gdt = st.adjoint_backward(gxN, gvN)

    opot = orc.milky_way() if use_pot else None
    states = ar.kdk_forward(pos, vel, m, eps2, st.ctx.timestep, steps, opot)
    rx, rv, rm, rdt = ar.kdk_adjoint(states, m, eps2, st.ctx.timestep,
                                     *_loss_grads(states[-1][0], states[-1][1], c0, wx, wv), opot)
    for name, got, ref in (("x", gx, rx), ("v", gv, rv), ("m", gm, rm)):
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(name, "max rel", err)
        assert err <= 1e-9, name
    assert abs(gdt[0] - rdt) <= 1e-9 * abs(rdt)
    # the state is left at the end of the forward run
    assert np.array_equal(st.bodies()[0], x)


def test_checkpoint_interval_does_not_change_gradients():
    pos, vel, m = _dwarf(2000, 3)
    pot = nb.milky_way_potential()
    rng = np.random.default_rng(1)
    wx, wv = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    out = []
    for every in (1, 4, 12):
        st = nb.NBodyState(_ctx(1e-3, 1e-3), pot, pos, vel, m, prestep=False)
        st.adjoint_forward(12, every)
        x, v, _ = st.bodies()
        out.append(st.adjoint_backward(*_loss_grads(x, v, pos.mean(0), wx, wv)))
        st.close()
    for o in out[1:]:
        for a, b in zip(o, out[0]):
            assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


@pytest.mark.parametrize("criterion,quad,use_pot", [("bh86", True, True), ("bh86", False, False),
                                                    ("sw93", True, True), ("treecode", True, False)])
def test_gravity_vjp_matches_frozen_tree_reference(criterion, quad, use_pot):
    """One force adjoint on the GPU (walk replay, node records summed across
    the wave, top-down hand-off, leaf scatter, potential Jacobian) against the
    numpy frozen-tree VJP on the oracle's identical tree and interaction
    lists (itself checked against central differences of the frozen-list
    force in tests/test_adjoint_ref.py)."""
    n, eps2 = 2000, 1e-3
    pos, vel, m = _dwarf(n, 5) if use_pot else plummer_bodies(n, seed=5)
    m = m * np.random.default_rng(7).uniform(0.5, 1.5, n)
    pot = nb.milky_way_potential() if use_pot else None
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=eps2, theta=0.6,
                             criterion=criterion, useQuad=quad)
    st = nb.NBodyState(ctx, pot, pos, vel, m)
    G = np.random.default_rng(8).normal(size=(n, 3))
    gx, gm = st.gravity_vjp(G)
    opot = orc.milky_way() if use_pot else None
    o = orc.Oracle(orc.make_ctx(eps2=eps2, theta=0.6, criterion=criterion, use_quad=quad,
                                potential=opot), pos, vel, m)
    assert o.gravity() == 0
    t = o.tree()
    ch = ar.tree_children(t)
    lists = ar.walk_lists(t)
    rx, rm = ar.frozen_vjp(pos, m, t, ch, lists, eps2, G, quad)
    if use_pot:
        rx = rx + ar.ext_jac(ar.pot_params(opot), pos, G)
    ex = np.abs(gx - rx).max() / np.abs(rx).max()
    em = np.abs(gm - rm).max() / np.abs(rm).max()
    print(criterion, quad, use_pot, "max rel: pos", ex, "mass", em)
    assert ex <= 1e-11 and em <= 1e-11


def test_tree_adjoint_at_theta_zero_equals_exact_adjoint():
    """theta = 0 opens every cell: the tree adjoint then reduces to the
    direct-sum adjoint (only the summation order differs)."""
    pos, vel, m = _dwarf(1500, 6)
    pot = nb.milky_way_potential()
    rng = np.random.default_rng(3)
    wx, wv = rng.normal(size=pos.shape), rng.normal(size=pos.shape)
    out = []
    for crit, theta in (("bh86", 0.0), ("exact", 0.0)):
        st = nb.NBodyState(_ctx(1e-3, 1e-3, crit, theta), pot, pos, vel, m, prestep=False)
        st.adjoint_forward(6, 4)
        x, v, _ = st.bodies()
        out.append(st.adjoint_backward(wx + (x - pos.mean(0)), wv))
        st.close()
    for name, a, b in zip(("x", "v", "m", "dt"), out[0], out[1]):
        err = np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()
        print(name, "tree(theta=0) vs exact max rel", err)
        assert err <= 1e-10, name


def test_batched_adjoint_equals_individual_runs():
    sizes = [700, 1500, 300]
    members = [_dwarf(k, 10 + i, offset=(8.0 + i, 3.0, -2.0 * i)) for i, k in enumerate(sizes)]
    pot = nb.milky_way_potential()
    ctx = _ctx(1e-3, 1e-3)
    rng = np.random.default_rng(4)
    ws = [(rng.normal(size=(k, 3)), rng.normal(size=(k, 3))) for k in sizes]
    bt = nb.NBodyBatch(ctx, pot, members, prestep=False)
    bt.adjoint_forward(9, 4)
    gxs, gvs = [], []
    for i, (p, _, _) in enumerate(members):
        x, v, _ = bt.sim_bodies(i)
        gx, gv = _loss_grads(x, v, p.mean(0), *ws[i])
        gxs.append(gx)
        gvs.append(gv)
    bx, bv, bm, bdt = bt.adjoint_backward(np.concatenate(gxs), np.concatenate(gvs))
    off = 0
    for i, (p, v0, m0) in enumerate(members):
        st = nb.NBodyState(ctx, pot, p, v0, m0, prestep=False)
        st.adjoint_forward(9, 4)
        x, v, _ = st.bodies()
        gx, gv, gm, gdt = st.adjoint_backward(*_loss_grads(x, v, p.mean(0), *ws[i]))
        k = sizes[i]
        for a, b in ((bx[off:off + k], gx), (bv[off:off + k], gv), (bm[off:off + k], gm)):
            assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), i
        assert abs(bdt[i] - gdt[0]) <= 1e-12 * abs(gdt[0]), i
        off += k
        st.close()


@pytest.mark.parametrize("use_pot", [False, True])
def test_tree_adjoint_over_steps_matches_frozen_reverse_sweep(use_pot):
    """The tree adjoint over several steps (theta = 0.6, quadrupoles) against
    the numpy reverse sweep over the oracle's trees of the same run
    (oracle/adjoint_ref.kdk_adjoint_frozen, pinned by central differences of
    the frozen-list forward map in tests/test_adjoint_ref.py): the GPU's
    trees are the oracle's node for node, so the two differentiate the same
    map and agree to round-off."""
    n, steps, eps2, dt = 1200, 8, 1e-4, 2e-4
    pos, vel, m = plummer_bodies(n, seed=33)
    m = m * np.random.default_rng(4).uniform(0.5, 1.5, n)
    if use_pot:
        pos = pos + np.array([8.0, 3.0, -2.0])
    pot = nb.milky_way_potential() if use_pot else None
    opot = orc.milky_way() if use_pot else None
    ctx = nb.NBodyCtx.create(timestep=dt, timeEvolve=dt * steps, eps2=eps2, theta=0.6)
    st = nb.NBodyState(ctx, pot, pos, vel, m, prestep=False)
    st.adjoint_forward(steps, 3)
    x, v, _ = st.bodies()
    rng = np.random.default_rng(9)
    c0 = pos.mean(0)
    wx, wv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    got = st.adjoint_backward(wx + (x - c0), wv)
    st.close()

    o = orc.Oracle(orc.make_ctx(eps2=eps2, theta=0.6, timestep=dt, potential=opot), pos, vel, m)
    assert o.gravity() == 0
    frozen = ar.frozen_trees(o, steps)
    assert np.array_equal(x, o.pos)
    states = ar.kdk_forward_frozen(pos, vel, m, eps2, dt, frozen, opot)
    want = ar.kdk_adjoint_frozen(states, m, eps2, dt, frozen, wx + (x - c0), wv, opot)
    for name, a, b in zip(("x", "v", "m", "dt"), got, want):
        a, b = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
        err = np.abs(a - b).max() / np.abs(b).max()
        print(name, "GPU tree adjoint vs frozen reverse sweep, max rel", err, flush=True)
        assert err <= 1e-10, name
