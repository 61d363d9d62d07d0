"""The benchmark configurations, each run through the HIP path on its own
workload:

* two-component dwarf, N = 10^5 (Plummer 12/0.2 + NFW 60/0.8 on its orbit
  in the Milky Way potential, dt = 1e-3): FP64 steps against the oracle bit
  for bit, and the 50 x 1 lambda-beta histogram of the baryons equal to the
  oracle's (EMD 0);
* the same model at N = 10^6 with the mixed walks: step-0 median and 99th
  percentile relative acceleration error against the FP64 direct sum on
  4096 sampled bodies, and against the FP64 tree body by body;
* 64 dwarfs drawn from dSFMT seed 7 (N ~ U(2000, 40000)) in one batch with
  hipGraph-replayed steps: every dwarf's histogram against its own single
  run, EMD <= 1e-10;
* the differentiation workload, N = 20 000, 1000 steps at p* + delta: the
  adjoint gradient of the six parameters against central differences
  (h = 1e-5 relative) to 1e-3, Exact criterion; and the Exact forward force
  at N >= 4096 (segmented sum) against the oracle to 1e-12.
"""
import numpy as np
import pytest
import torch

from helpers import rel_err
from milkywayathome_client_amd import dwarf, histogram as H, models, nbody as nb
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _dwarf(n, seed=34086709):
    pot = nb.milky_way_potential()
    x0 = models.lbr_to_cartesian([218.0, 53.5, 28.6])
    fpos, fvel = models.reverse_orbit(pot, x0, [-156.0, 79.0, 107.0], 4.0, 1e-4)
    pos, vel, m, _ = dwarf.two_component_dwarf(n, seed=seed, position=fpos, velocity=fvel)
    eps2 = models.calculate_eps2(n, 0.2, 0.8, 12.0, 60.0)
    return pot, pos, vel, m, eps2


def test_dwarf_100k_steps_and_histogram_match_oracle():
    n, k = 100000, 8
    pot, pos, vel, m, eps2 = _dwarf(n)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=4.0, eps2=eps2, theta=0.5)
    print(f"N={n} dwarf, {k} FP64 graph steps against the oracle", flush=True)
    gpu = nb.NBodyState(ctx, pot, pos, vel, m)
    gpu.set_graph(True)
    o = orc.Oracle(orc.make_ctx(eps2=eps2, theta=0.5, timestep=ctx.timestep,
                                potential=orc.milky_way()), pos, vel, m, threads=16)
    assert o.gravity() == 0
    assert gpu.steps(k) == 0
    for _ in range(k):
        assert o.step() == 0
    p, v, a = gpu.bodies()
    assert np.array_equal(p, o.pos) and np.array_equal(v, o.vel) and np.array_equal(a, o.acc)
    # the workunit's histogram of the baryons (the dark matter is ignored)
    hp = H.HistogramParams()
    nbar = n // 2
    gpu.set_ignore(np.arange(n) >= nbar)
    counts, totals = gpu.histogram(hp)
    frac_o, tot_o = H.create_histogram(o.pos[:nbar], hp)
    assert int(totals[0]) == tot_o
    frac_g = counts[0].astype(np.float64) / max(int(totals[0]), 1)
    assert np.array_equal(frac_g, frac_o)
    assert H.emd(frac_g, frac_o, hp) == 0.0
    gpu.close()


@pytest.fixture(scope="module")
def million():
    pot, pos, vel, m, eps2 = _dwarf(1000000)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=eps2, theta=0.5)
    st = nb.NBodyState(ctx, pot, pos, vel, m)
    rng = np.random.default_rng(1)
    idx = np.sort(rng.choice(pos.shape[0], size=4096, replace=False)).astype(np.int32)
    direct = st.direct_accel(idx) + models.ext_accel(pot, pos[idx])
    a64 = st.accelerations()
    st.close()
    return ctx, pot, pos, vel, m, idx, direct, a64


def _total_err(a, direct, idx):
    return np.linalg.norm(a[idx] - direct, axis=1) / np.linalg.norm(direct, axis=1)


@pytest.mark.parametrize("precision", ["mixed", "mixed_coherent", "mixed_group"])
def test_million_body_mixed_walks_at_step_0(million, precision):
    ctx, pot, pos, vel, m, idx, direct, a64 = million
    st = nb.NBodyState(ctx, pot, pos, vel, m, precision=precision)
    a = st.accelerations()
    e_direct = _total_err(a, direct, idx)
    e_fp64 = _total_err(a64, direct, idx)
    e_tree = rel_err(a, a64, 0)
    print(f"N=1e6 {precision}: vs direct median {np.median(e_direct):.3e} "
          f"p99 {np.percentile(e_direct, 99):.3e} (FP64 tree {np.median(e_fp64):.3e}); "
          f"vs FP64 tree median {np.median(e_tree):.3e} max {e_tree.max():.3e}", flush=True)
    assert st.tree_info().status == 0
    if precision == "mixed":
        # the reference lists in FP32: the FP64 tree's error, nothing added
        assert np.median(e_tree) <= 1e-7 and np.percentile(e_tree, 99) <= 1e-6
        assert abs(np.median(e_direct) - np.median(e_fp64)) <= 1e-7
    else:
        # lists that refine the reference lists: at most the FP64 tree's error
        assert np.median(e_direct) <= np.median(e_fp64)
        assert np.percentile(e_direct, 99) <= 1.05 * np.percentile(e_fp64, 99)
    st.close()


def test_64_dwarfs_batched_histograms_equal_single_runs():
    import bench
    pot, members, eps2, params = bench.make_batch_workload(64)
    assert len(members) == 64
    k = 12
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=2.0, eps2=1e-4, theta=0.5)
    print(f"64 dwarfs, {sum(p[0] for p in params)} bodies, {k} graph steps", flush=True)
    bt = nb.NBodyBatch(ctx, pot, members, eps2=eps2)
    bt.set_graph(True)
    assert bt.steps(k) == 0
    assert bt.statuses() == [0] * 64
    hp = H.HistogramParams()
    ignore = np.concatenate([np.arange(p[0]) >= p[0] // 2 for p in params])
    bt.set_ignore(ignore)
    counts, totals = bt.histogram(hp)
    worst = 0.0
    for s, (pos, vel, m) in enumerate(members):
        ck = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=2.0, eps2=eps2[s], theta=0.5)
        st = nb.NBodyState(ck, pot, pos, vel, m)
        assert st.steps(k) == 0
        st.set_ignore(np.arange(pos.shape[0]) >= params[s][0] // 2)
        c1, t1 = st.histogram(hp)
        h1 = c1[0].astype(np.float64) / max(int(t1[0]), 1)
        hb = counts[s].astype(np.float64) / max(int(totals[s]), 1)
        e = H.emd(hb, h1, hp)
        worst = max(worst, e)
        assert e <= 1e-10, (s, e)
        assert np.array_equal(st.bodies()[0], bt.sim_bodies(s)[0]), s
        st.close()
    print("largest EMD batched vs single", worst, flush=True)
    bt.close()


def test_exact_forward_beyond_one_segment_matches_oracle():
    """N = 8192: the Exact sum is split into segments (one per grid row)
    above 4096 bodies, so its order differs from the reference's serial
    sum; the difference stays at rounding level."""
    n = 8192
    pot, pos, vel, m, _ = _dwarf(n, seed=77)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-3, criterion="Exact")
    st = nb.NBodyState(ctx, pot, pos, vel, m)
    o = orc.Oracle(orc.make_ctx(eps2=1e-3, theta=0.0, timestep=1e-3, criterion="exact",
                                potential=orc.milky_way()), pos, vel, m, threads=16)
    assert o.gravity() == 0
    e = rel_err(st.accelerations(), o.acc, 0)
    print(f"Exact N={n}: max rel diff {e.max():.3e}", flush=True)
    assert e.max() <= 1e-12
    assert st.steps(3) == 0
    for _ in range(3):
        assert o.step() == 0
    p, _, _ = st.bodies()
    np.testing.assert_allclose(p, o.pos, rtol=1e-12, atol=1e-12)
    st.close()


def test_differentiation_workload_gradient_matches_central_differences():
    """20 000 bodies, 1000 steps (forward time 1 Gyr at dt = 1e-3),
    checkpoints every 50 steps, Exact criterion, p* + N(0, 0.02 |p*|)."""
    from milkywayathome_client_amd import adjoint as adj
    p_star = np.array([1.0, 1.0, 0.2, 0.2, 12.0, 0.2])
    model = adj.make_model(n=20000, p_ref=p_star, criterion="exact", every=50)
    assert model.cfg.nsteps == 1000
    with torch.no_grad():
        target = model.histogram(torch.tensor(p_star), grad=False)
    rng = np.random.default_rng(11)
    p = p_star + rng.normal(size=6) * 0.02 * np.abs(p_star)
    print("differentiation workload: gradient and 12 forward runs", flush=True)
    L, g = model.value_and_grad(p, target)
    fd = np.zeros(6)
    for k in range(6):
        h = 1e-5 * abs(p[k])
        a, b = p.copy(), p.copy()
        a[k] += h
        b[k] -= h
        with torch.no_grad():
            fd[k] = (float(model.loss(torch.tensor(a), target, grad=False))
                     - float(model.loss(torch.tensor(b), target, grad=False))) / (2 * h)
    rel = np.abs(g - fd) / np.abs(fd)
    print("loss", L, "\nadjoint", g, "\ncentral", fd, "\nrel", rel, flush=True)
    assert L > 0 and np.all(np.isfinite(g))
    assert rel.max() <= 1e-3
