"""Two-component dwarf (the second workload's shape: Plummer 12/0.2 + NFW
60/0.8, on its orbit in the Milky Way potential, dt = 1e-3) through the
plain FP64 path against the oracle, plus size-independent properties at the
full 10^5 and 10^6 sizes."""
import numpy as np
import pytest

from helpers import rel_err
from milkywayathome_client_amd import dwarf, models, nbody as nb
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _dwarf(n, seed=34086709):
    pot = nb.milky_way_potential()
    x0 = models.lbr_to_cartesian([218.0, 53.5, 28.6])
    fpos, fvel = models.reverse_orbit(pot, x0, [-156.0, 79.0, 107.0], 4.0, 1e-4)
    pos, vel, m, _ = dwarf.two_component_dwarf(n, seed=seed, position=fpos, velocity=fvel)
    eps2 = models.calculate_eps2(n, 0.2, 0.8, 12.0, 60.0)
    return pot, pos, vel, m, eps2


def test_dwarf_100k_matches_oracle():
    pot, pos, vel, m, eps2 = _dwarf(100000)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=4.0, eps2=eps2, theta=0.5)
    gpu = nb.NBodyState(ctx, pot, pos, vel, m)
    o = orc.Oracle(orc.make_ctx(eps2=eps2, theta=0.5, timestep=ctx.timestep,
                                potential=orc.milky_way()), pos, vel, m, threads=16)
    assert o.gravity() == 0
    assert gpu.tree_info().nnode == o.tree_summary()[2]
    assert rel_err(gpu.accelerations(), o.acc).max() <= 1e-12
    for _ in range(3):
        assert gpu.step() == 0 and o.step() == 0
    p, v, a = gpu.bodies()
    np.testing.assert_allclose(p, o.pos, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(v, o.vel, rtol=1e-12, atol=1e-9)


def test_theta_zero_equals_direct_sum_at_scale():
    """Size-independent identity at N = 10^5: theta = 0 walk == direct sum."""
    pot, pos, vel, m, eps2 = _dwarf(100000, seed=5)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=eps2, theta=0.0)
    st = nb.NBodyState(ctx, None, pos, vel, m)
    idx = np.arange(0, 100000, 97, dtype=np.int32)
    assert rel_err(st.accelerations()[idx], st.direct_accel(idx), 0).max() < 1e-11


def test_million_body_step_is_consistent():
    """N = 10^6 (the benchmark size): the tree holds every massive body
    once, Next links partition the table, and the momentum carried by
    self-gravity sums to ~0 on a sample against the direct sum."""
    pot, pos, vel, m, eps2 = _dwarf(1000000)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=eps2, theta=0.5)
    st = nb.NBodyState(ctx, pot, pos, vel, m)
    info = st.tree_info()
    assert info.nmassive == 1000000 and info.status == 0
    t = st.tree()
    assert np.array_equal(np.sort(t["ref"][t["kind"] == 1]), np.arange(1000000))
    assert np.all(t["skip"] > np.arange(info.nnode))
    assert t["skip"][0] == info.nnode
    np.testing.assert_allclose(t["mass"][0], m.sum(), rtol=1e-12)
    np.testing.assert_allclose(t["pos"][0], (m[:, None] * pos).sum(0) / m.sum(), rtol=1e-12)
    idx = np.arange(0, 1000000, 997, dtype=np.int32)
    d = st.direct_accel(idx)
    self_tree = st.walk_accel()[idx]
    e = rel_err(self_tree, d, 0)
    print("1M self-gravity tree vs direct: median", np.median(e))
    assert np.median(e) < 2e-3
    assert st.step() == 0
