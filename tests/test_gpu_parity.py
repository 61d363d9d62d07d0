"""Parity of the HIP plain (FP64) path with the CPU oracle, through the C ABI.

The oracle restates the reference CPU tree code (nbody/src/nbody_tree.c,
nbody_grav.c, nbody_plain.c); the plain form must match it to 1e-12 relative
per body (in practice bit for bit: both keep the reference operation order).
"""
import numpy as np
import pytest

from helpers import plummer_bodies, rel_err
from milkywayathome_client_amd import nbody as nb
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _pair(n, theta=0.5, criterion="bh86", quad=True, potential=False, seed=42, eps2=1e-4,
          dt=1e-4, bodies=None, rsize=4.0, allow_incest=False):
    pos, vel, m = bodies if bodies is not None else plummer_bodies(n, seed=seed)
    pot = nb.milky_way_potential() if potential else None
    ctx = nb.NBodyCtx.create(timestep=dt, timeEvolve=dt * 200, eps2=eps2, theta=theta,
                             criterion=criterion, useQuad=quad, treeRSize=rsize,
                             allowIncest=allow_incest)
    gpu = nb.NBodyState(ctx, pot, pos, vel, m)
    octx = orc.make_ctx(eps2=eps2, theta=ctx.theta, timestep=ctx.timestep, criterion=criterion,
                        use_quad=ctx.useQuad, tree_rsize=rsize, allow_incest=allow_incest,
                        potential=orc.milky_way() if potential else None)
    cpu = orc.Oracle(octx, pos, vel, m)
    return gpu, cpu


def test_tree_matches_oracle_node_for_node():
    gpu, cpu = _pair(3000)
    assert cpu.gravity() == 0
    gt, ct = gpu.tree(), cpu.tree()
    info = gpu.tree_info()
    ncell, maxdepth, nnode = cpu.tree_summary()
    print("nnode", nnode, "ncell", ncell, "maxdepth", maxdepth, "rsize", cpu.rsize)
    assert info.nnode == nnode and info.ncell == ncell and info.max_depth == maxdepth
    assert info.rsize == cpu.rsize
    for k in ("kind", "skip", "level"):
        assert np.array_equal(gt[k], ct[k]), k
    bodies = ct["kind"] == 1
    assert np.array_equal(gt["ref"][bodies], ct["ref"][bodies])
    for k in ("pos", "mass", "rcrit2", "quad"):
        np.testing.assert_array_equal(gt[k], ct[k], err_msg=k)


@pytest.mark.parametrize("criterion,theta,quad", [
    ("bh86", 0.5, True), ("bh86", 0.7, False), ("sw93", 0.6, True), ("treecode", 0.5, True),
    ("bh86", 1.0, True)])
def test_accelerations_match_oracle(criterion, theta, quad):
    gpu, cpu = _pair(4000, theta=theta, criterion=criterion, quad=quad, allow_incest=True)
    rc = cpu.gravity()
    a_gpu = gpu.accelerations()
    err = rel_err(a_gpu, cpu.acc)
    print(criterion, theta, quad, "max rel", err.max(), "bitwise", np.mean(a_gpu == cpu.acc))
    assert err.max() <= TOL
    assert (gpu.last_status != 0) == (rc != 0)


def test_theta_zero_walk_equals_direct_sum():
    gpu, cpu = _pair(2000, theta=0.0)
    pos, _, m = plummer_bodies(2000)
    d = orc.direct_accel(pos, m, 1e-4)
    err = rel_err(gpu.accelerations(), d)
    print("theta=0 vs direct", err.max())
    assert err.max() < 1e-12


def test_gpu_direct_sum_checker():
    gpu, _ = _pair(3000)
    pos, _, m = plummer_bodies(3000)
    idx = np.arange(0, 3000, 7, dtype=np.int32)
    d = orc.direct_accel(pos, m, 1e-4, idx)
    g = gpu.direct_accel(idx)
    assert rel_err(g, d).max() < 1e-12


def test_external_potential_and_steps_match_oracle():
    pos, vel, m = plummer_bodies(3000, position=(5.0, -3.0, 2.0), velocity=(10.0, 50.0, -20.0))
    gpu, cpu = _pair(3000, potential=True, bodies=(pos, vel, m), dt=1e-3)
    assert cpu.gravity() == 0
    assert rel_err(gpu.accelerations(), cpu.acc).max() <= TOL
    for _ in range(10):
        assert gpu.step() == 0
        assert cpu.step() == 0
    p, v, a = gpu.bodies()
    print("pos max abs diff", np.abs(p - cpu.pos).max(), "vel", np.abs(v - cpu.vel).max())
    np.testing.assert_allclose(p, cpu.pos, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(v, cpu.vel, rtol=1e-12, atol=1e-10)
    assert rel_err(a, cpu.acc).max() <= TOL
    assert gpu.tree_info().rsize == cpu.rsize


def test_exact_criterion_matches_oracle():
    gpu, cpu = _pair(1500, criterion="exact", potential=True)
    assert cpu.gravity() == 0
    assert rel_err(gpu.accelerations(), cpu.acc).max() <= TOL


def test_massless_bodies_feel_force_but_are_not_in_tree():
    pos, vel, m = plummer_bodies(1000)
    m = m.copy()
    m[::10] = 0.0
    gpu, cpu = _pair(1000, bodies=(pos, vel, m), allow_incest=True)
    rc = cpu.gravity()
    info = gpu.tree_info()
    assert info.nmassive == int(np.count_nonzero(m))
    assert info.nnode == cpu.tree_summary()[2]
    # a massless body never meets itself in the walk: the reference reports incest
    assert rc == nb.NBODY_TREE_INCEST_NONFATAL
    assert gpu.last_status == nb.NBODY_TREE_INCEST_NONFATAL
    assert rel_err(gpu.accelerations(), cpu.acc).max() <= TOL


@pytest.mark.parametrize("n", [1, 2, 3, 17])
def test_tiny_systems(n):
    pos, vel, m = plummer_bodies(max(n, 2))
    pos, vel, m = pos[:n], vel[:n], m[:n]
    gpu, cpu = _pair(n, bodies=(pos, vel, m))
    cpu.gravity()
    assert gpu.tree_info().nnode == cpu.tree_summary()[2]
    assert rel_err(gpu.accelerations(), cpu.acc).max() <= TOL if n > 1 else True


def test_rsize_grows_and_persists():
    pos, vel, m = plummer_bodies(500)
    pos = pos.copy()
    pos[0] = (40.0, -3.0, 1.0)
    gpu, cpu = _pair(500, bodies=(pos, vel, m))
    cpu.gravity()
    assert gpu.tree_info().rsize == cpu.rsize == 128.0


def test_plummer_cfg1_200_steps():
    """Plummer N=10,000, seed 42, M=12, a=0.2, no external potential, BH86
    theta=0.5 with quadrupoles, eps2=1e-4, dt=1e-4 for 200 steps."""
    pos, vel, m = plummer_bodies(10000, seed=42)
    gpu, cpu = _pair(10000, bodies=(pos, vel, m), dt=1e-4)
    assert cpu.gravity() == 0
    assert rel_err(gpu.accelerations(), cpu.acc).max() <= TOL
    assert gpu.steps(200) == 0
    for _ in range(200):
        cpu.step()
    p, v, a = gpu.bodies()
    dp = np.abs(p - cpu.pos).max()
    print("cfg1 final pos max abs diff", dp, "bitwise frac", np.mean(p == cpu.pos))
    np.testing.assert_allclose(p, cpu.pos, rtol=0, atol=1e-12)
    assert rel_err(a, cpu.acc).max() <= 1e-10
