"""GPU lambda-beta histogram (nbhip_histogram) against the host restatement
of nbCreateHistogram's binning (milkywayathome_client_amd/histogram.py,
itself pinned by tests/test_histogram.py): identical counts and totals,
for one state and per simulation of a batch, with ignored bodies left out.
Bodies within 1e-9 degrees of a bin edge could fall either side (device and
host atan2/asin may differ in the last bit); the workloads below have none."""
import numpy as np
import pytest

from helpers import plummer_bodies
from milkywayathome_client_amd import histogram as H
from milkywayathome_client_amd import nbody as nb

pytestmark = pytest.mark.gpu


def _stream(n, seed, centre):
    rng = np.random.default_rng(seed)
    # bodies spread along a ring through the frame so many bins are filled
    t = rng.uniform(0, 2 * np.pi, n)
    pos = np.stack([30 * np.cos(t), 30 * np.sin(t), rng.normal(0, 3, n)], 1) + centre
    vel = rng.normal(0, 10, (n, 3))
    return pos, vel, np.full(n, 1e-3)


def _edges_clear(pos, hp):
    lam, beta = H.lambda_beta(pos, hp)
    w = (hp.lambdaEnd - hp.lambdaStart) / hp.lambdaBins
    fl = (lam - hp.lambdaStart) / w
    return np.all(np.abs(fl - np.round(fl)) > 1e-9)


@pytest.mark.parametrize("beta_bins", [1, 3])
def test_histogram_matches_host(beta_bins):
    hp = H.HistogramParams(betaBins=beta_bins)
    pos, vel, m = _stream(20000, 1, (0.0, 0.0, 0.0))
    ignore = np.zeros(len(m), dtype=bool)
    ignore[::3] = True
    assert _edges_clear(pos, hp)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-3, theta=0.5)
    st = nb.NBodyState(ctx, None, pos, vel, m, prestep=False)
    st.set_ignore(ignore)
    counts, totals = st.histogram(hp)
    ref, tot = H.create_histogram(pos, hp, ignore=ignore)
    assert totals[0] == tot
    np.testing.assert_array_equal(counts[0], np.round(ref * tot).astype(np.int64))
    st.set_ignore(None)
    counts2, totals2 = st.histogram(hp)
    ref2, tot2 = H.create_histogram(pos, hp)
    assert totals2[0] == tot2 and totals2[0] > totals[0]


def test_batch_histograms_per_simulation():
    hp = H.HistogramParams()
    members = [_stream(k, 10 + i, (5.0 * i, 0.0, 0.0)) for i, k in enumerate([900, 7000, 1, 4000])]
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-3, theta=0.5)
    bt = nb.NBodyBatch(ctx, None, members, prestep=False)
    ignore = np.concatenate([np.arange(len(mm[2])) % 2 == 1 for mm in members])
    bt.set_ignore(ignore)
    counts, totals = bt.histogram(hp)
    off = 0
    for k, (pos, _, m) in enumerate(members):
        ig = ignore[off:off + len(m)]
        ref, tot = H.create_histogram(pos, hp, ignore=ig)
        assert totals[k] == tot, k
        np.testing.assert_array_equal(counts[k], np.round(ref * tot).astype(np.int64))
        off += len(m)


def test_likelihood_from_gpu_counts():
    """nbSystemLikelihood on the GPU histogram (likelihood.from_counts) equals
    the one on the host binning of the same bodies; a run against its own
    histogram scores 0, and against another run a positive geometry term."""
    from milkywayathome_client_amd import likelihood as L
    hp = H.HistogramParams()
    pos, vel, m = _stream(20000, 3, (0.0, 0.0, 0.0))
    pos2, _, _ = _stream(20000, 4, (2.0, 0.0, 0.0))
    assert _edges_clear(pos, hp) and _edges_clear(pos2, hp)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-3, theta=0.5)
    st = nb.NBodyState(ctx, None, pos, vel, m, prestep=False)
    counts, totals = st.histogram(hp)
    sim = L.from_counts(counts[0], hp, total_simulated=len(m), mass_per_particle=1e-3)
    ref, tot = H.create_histogram(pos, hp)
    host = L.from_counts(np.round(ref * tot).astype(np.int64), hp, total_simulated=len(m),
                         mass_per_particle=1e-3)
    np.testing.assert_array_equal(sim.variable, host.variable)
    assert np.all(L.system_likelihood(host, sim, len(m))[:3] == 0.0)
    st2 = nb.NBodyState(ctx, None, pos2, vel, m, prestep=False)
    other = L.from_counts(st2.histogram(hp)[0][0], hp, total_simulated=len(m),
                          mass_per_particle=1e-3)
    out = L.system_likelihood(other, sim, len(m))
    assert out[1] > 0.0 and np.isfinite(out[0])
    st.close()
    st2.close()


def test_dispersions_match_host_restatement():
    """nbhip_histogram_dispersion (histograms 1-7: beta and line-of-sight
    velocity dispersions, average line-of-sight velocity, beta, distance and
    proper motions, with outlier rejection) against the numpy restatement of the reference's
    loop (oracle/histogram_ref.py): same bins, same rejections, the sums
    differing only in association (1e-12); for one state with ignored bodies
    and for each member of a batch."""
    from oracle import histogram_ref as R
    hp = H.HistogramParams()
    pos, vel, m = _stream(40000, 7, (0.0, 0.0, 0.0))
    rng = np.random.default_rng(8)
    vel[rng.choice(len(m), 50, replace=False)] *= 20.0        # velocity outliers
    ignore = np.zeros(len(m), dtype=bool)
    ignore[::5] = True
    assert _edges_clear(pos, hp)
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-3, theta=0.5)
    st = nb.NBodyState(ctx, None, pos, vel, m, prestep=False)
    st.set_ignore(ignore)
    got = st.histogram_dispersion(hp)
    ref = R.dispersions(pos, vel, hp, ignore=ignore)
    for k in range(1, 8):
        for g, r in zip(got[k], ref[k]):
            np.testing.assert_allclose(g, r, rtol=1e-12, atol=1e-12)
    assert np.count_nonzero(got[1][0]) >= 15 and np.count_nonzero(got[2][0]) >= 15
    assert np.count_nonzero(got[5][1] > 0) >= 15
    assert np.count_nonzero(got[6][1] > 0) >= 15 and np.count_nonzero(got[7][1] > 0) >= 15
    st.close()
    members = [_stream(k, 20 + i, (4.0 * i, 0.0, 0.0)) for i, k in enumerate([3000, 25000, 8000])]
    bt = nb.NBodyBatch(ctx, None, members, prestep=False)
    for k, (p, v, _) in enumerate(members):
        got = bt.histogram_dispersion(hp, sim=k)
        ref = R.dispersions(p, v, hp)
        for j in range(1, 8):
            for g, r in zip(got[j], ref[j]):
                np.testing.assert_allclose(g, r, rtol=1e-12, atol=1e-12)
    bt.close()


def test_dispersion_scratch_reuse_across_shapes():
    """The dispersion scratch is kept by the state and grown on demand: calls
    with a small, a larger and again the small histogram shape on one state
    each match the restatement."""
    from oracle import histogram_ref as R
    pos, vel, m = _stream(20000, 11, (0.0, 0.0, 0.0))
    ctx = nb.NBodyCtx.create(timestep=1e-3, timeEvolve=1.0, eps2=1e-3, theta=0.5)
    st = nb.NBodyState(ctx, None, pos, vel, m, prestep=False)
    shapes = [H.HistogramParams(), H.HistogramParams(lambdaBins=100, betaBins=3),
              H.HistogramParams()]
    for hp in shapes:
        assert _edges_clear(pos, hp)
        got = st.histogram_dispersion(hp)
        ref = R.dispersions(pos, vel, hp)
        for k in range(1, 8):
            for g, r in zip(got[k], ref[k]):
                np.testing.assert_allclose(g, r, rtol=1e-12, atol=1e-12)
    st.close()
