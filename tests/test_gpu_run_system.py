"""nbhip_run_system, the C run loop with nbRunSystemPlain's hooks
(nbody/src/nbody_plain.c:366-461), through the C ABI:

* the best likelihood it keeps over graph-replayed steps equals the one the
  oracle's run gives (the oracle stepped on the CPU, its bodies binned on
  the host, the same likelihood) and the one of the step-by-step Python
  loop, with the same step, time and number of improvements, and its kept
  bodies are the oracle's at that step;
* the likelihood hook runs after every step from BestLikeStart on, the
  checkpoint hook after every chunk before it and every step after it;
* a fatal status stops the run at the step that failed, also with graph
  steps (later replays do nothing and are not counted).
"""
import numpy as np
import pytest

from milkywayathome_client_amd import histogram as H
from milkywayathome_client_amd import likelihood as L
from milkywayathome_client_amd import nbody as nb
from milkywayathome_client_amd import plain as P
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MASS = 1e-3


def _rotating_ring(n, seed, vscale=1.0):
    rng = np.random.default_rng(seed)
    t = rng.uniform(0, 2 * np.pi, n)
    r = rng.uniform(26.0, 34.0, n)
    pos = np.stack([r * np.cos(t), r * np.sin(t), rng.normal(0, 2, n)], 1)
    vel = vscale * np.stack([-200.0 * np.sin(t), 200.0 * np.cos(t), np.zeros(n)], 1)
    vel += rng.normal(0, 5, (n, 3))
    return pos, vel, np.full(n, MASS)


def _likelihood_of_counts(data, counts, hp, n):
    hist = L.from_counts(counts, hp, n, MASS)
    return np.asarray(L.system_likelihood(data, hist, n), dtype=np.float64)


def test_run_system_best_likelihood_matches_oracle_and_python_loop():
    hp = H.HistogramParams()
    n = 5000
    pos, vel, m = _rotating_ring(n, 3)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=0.02, eps2=1e-3, theta=0.5)
    assert ctx.nStep == 20
    # data: a slightly faster ring at step 14, so the likelihood has a clear
    # minimum that is not 0
    dpos, dvel, dm = _rotating_ring(n, 3, vscale=1.02)
    twin = nb.NBodyState(ctx, None, dpos, dvel, dm)
    twin.run(14)
    data = P.simulated_histogram(twin, hp, n, MASS, dispersions=False)
    twin.close()

    st = nb.NBodyState(ctx, None, pos, vel, m, prestep=False)
    st.set_graph(True)
    seen_like, seen_ckpt = [], []

    def like(step):
        seen_like.append(step)
        counts, _ = st.histogram(hp)
        return _likelihood_of_counts(data, counts[0], hp, n)

    def ckpt(step):
        seen_ckpt.append(step)
        return 0

    print("run_system: 20 steps, BestLikeStart 0.5, graph chunks of 4", flush=True)
    rc, best = st.run_system(like, ckpt, best_like_start=0.5, keep_best_bodies=True, chunk=4)
    assert rc == 0 and st.tree_info().step == 20
    assert seen_like == list(range(10, 21))
    assert seen_ckpt == [4, 8, 9] + list(range(10, 21))
    assert best["evaluated"] == 11

    # the Python loop over single eager steps
    st2 = nb.NBodyState(ctx, None, pos, vel, m, prestep=False)
    ref = P.run_best_likelihood(st2, data, hp, n, MASS, best_like_start=0.5, record=True)
    st2.close()
    # the oracle: CPU steps, host binning
    o = orc.Oracle(orc.make_ctx(eps2=1e-3, theta=0.5, timestep=ctx.timestep), pos, vel, m)
    assert o.gravity() == 0
    ob = {"like": L.DEFAULT_WORST_CASE, "step": 0, "count": 0, "pos": None}
    for step in range(1, 21):
        assert o.step() == 0
        if step / 20.0 < 0.5:
            continue
        frac, total = H.create_histogram(o.pos, hp)
        counts = np.rint(frac * total).astype(np.uint32)
        lk = P.clamp_likelihood(float(_likelihood_of_counts(data, counts, hp, n)[0]))
        if abs(lk) < abs(ob["like"]):
            ob.update(like=lk, step=step, count=ob["count"] + 1, pos=o.pos)
    print("best: run_system", best["lk"][0], best["step"], best["count"], "| python",
          ref.likelihood, ref.step, ref.count, "| oracle", ob["like"], ob["step"], ob["count"],
          flush=True)
    assert best["lk"][0] == ref.likelihood == ob["like"]
    assert best["step"] == ref.step == ob["step"] and best["count"] == ref.count == ob["count"]
    assert best["time"] == ref.time == best["step"] / 20.0 * ctx.timeEvolve
    assert best["lk"][0] > P.DEFAULT_BEST_CASE
    bpos, _ = st.best_bodies()
    assert np.array_equal(bpos, ob["pos"])
    st.close()


def test_run_system_without_best_likelihood_equals_run():
    pos, vel, m = _rotating_ring(3000, 5)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=0.012, eps2=1e-3, theta=0.5)
    a = nb.NBodyState(ctx, None, pos, vel, m, prestep=False)
    a.set_graph(True)
    rc, best = a.run_system(use_best_like=False, chunk=5)
    assert rc == 0 and best["evaluated"] == 0 and best["lk"][0] == L.DEFAULT_WORST_CASE
    b = nb.NBodyState(ctx, None, pos, vel, m, prestep=False)
    b.run()
    for x, y in zip(a.bodies(), b.bodies()):
        assert np.array_equal(x, y)
    assert a.tree_info().step == b.tree_info().step == ctx.nStep
    with pytest.raises(nb.NBodyError):
        a.best_bodies()                 # nothing was kept
    a.close()
    b.close()


@pytest.mark.parametrize("graph", [False, True])
def test_fatal_status_stops_the_steps(graph):
    """A body whose velocity is NaN drifts to a NaN position in step 1: the
    root-size update reports NBODY_ERROR and the run stops after that step."""
    pos, vel, m = _rotating_ring(2000, 7)
    vel[17] = np.nan
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=0.01, eps2=1e-3, theta=0.5)
    st = nb.NBodyState(ctx, None, pos, vel, m)
    st.set_graph(graph)
    print("NaN body, graph", graph, flush=True)
    rc = st.lib.nbhip_steps(st._h, 6)
    assert nb.status_is_fatal(rc) and rc & nb.NBODY_ERROR
    assert st.tree_info().step == 1
    # nothing more runs
    rc = st.lib.nbhip_steps(st._h, 3)
    assert nb.status_is_fatal(rc) and st.tree_info().step == 1
    st.close()
    st = nb.NBodyState(ctx, None, pos, vel, m, prestep=False)
    st.set_graph(graph)
    calls = []
    with pytest.raises(nb.NBodyError):
        st.run_system(lambda s: calls.append(s) or [1.0] * 10, best_like_start=0.0, chunk=4)
    assert st.tree_info().step == 1 and calls == [1]     # evaluated, then stopped
    st.close()


def test_likelihood_hook_failure_stops_the_run():
    pos, vel, m = _rotating_ring(2000, 9)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=0.01, eps2=1e-3, theta=0.5)
    st = nb.NBodyState(ctx, None, pos, vel, m, prestep=False)

    def like(step):
        if step == 7:
            raise ValueError("no data histogram")
        return [float(step)] + [0.0] * 9

    with pytest.raises(ValueError):
        st.run_system(like, best_like_start=0.5)
    assert st.tree_info().step == 7
    st.close()
